/*
 * Bessel and Hankel functions of orders 0 and 1 in FP64 (oracle, test-only).
 *
 * The reference evaluates the Green's function with libm j0/y0
 * (B/MKL_GMRES_test_09122019/source/loop_functions.h:141-142) and with
 * scipy.special.hankel1 (B/PyGMRES/GMRES.py:20-21).  This is an independent
 * restatement from published mathematics:
 *   - x < 20: Miller's backward recurrence for J_n with the normalisation
 *     J0 + 2*sum J_2k = 1 (Abramowitz & Stegun 9.1.87 / 9.12), Neumann series
 *     Y0 = (2/pi)[(ln(x/2)+gamma) J0 - 2 sum (-1)^k J_2k / k]   (A&S 9.1.88),
 *     Y1 = -Y0' differentiated term by term with J_n' = (J_{n-1}-J_{n+1})/2.
 *   - x >= 20: Hankel's asymptotic expansion (DLMF 10.17.5),
 *     H_nu(x) ~ sqrt(2/(pi x)) e^{i(x - nu pi/2 - pi/4)} sum_k i^k a_k(nu)/x^k,
 *     a_k = a_{k-1} (4nu^2 - (2k-1)^2)/(8k), summed to its smallest term.
 * Pinned by tests/test_oracle.py against scipy.special values committed in
 * tests/golden/hankel_scipy.json and by the Wronskian J1 Y0 - J0 Y1 = 2/(pi x).
 */
#include <math.h>
#include "orc.h"

#define ORC_PI 3.14159265358979323846
#define ORC_EULER 0.57721566490153286061

static void bessel_miller(double x, double *J0, double *J1, double *Y0, double *Y1) {
  /* Backward recurrence J_{n-1} = (2n/x) J_n - J_{n+1} from a start order far
   * above x, with rescaling to stay in range; then normalise. */
  enum { NMAX = 160 };
  double J[NMAX + 2];
  int n0 = (int)(x + 12.0 * cbrt(x) + 40.0);
  if (n0 & 1) n0++;
  if (n0 > NMAX) n0 = NMAX;
  J[n0 + 1] = 0.0;
  J[n0] = 1e-30;
  for (int n = n0; n >= 1; --n) {
    J[n - 1] = (2.0 * n / x) * J[n] - J[n + 1];
    if (fabs(J[n - 1]) > 1e250) {
      for (int m = n - 1; m <= n0 + 1; ++m) J[m] *= 1e-250;
    }
  }
  double s_norm = 0.0, s_y0 = 0.0, s_y1 = 0.0;
  for (int k = 1; 2 * k <= n0 - 1; ++k) {
    const double sg = (k & 1) ? -1.0 : 1.0;
    s_norm += J[2 * k];
    s_y0 += sg * J[2 * k] / k;
    s_y1 += sg * (J[2 * k - 1] - J[2 * k + 1]) / k;
  }
  const double norm = J[0] + 2.0 * s_norm;
  const double j0 = J[0] / norm, j1 = J[1] / norm;
  s_y0 /= norm;
  s_y1 /= norm;
  const double lg = log(0.5 * x) + ORC_EULER;
  *J0 = j0;
  *J1 = j1;
  *Y0 = (2.0 / ORC_PI) * (lg * j0 - 2.0 * s_y0);
  *Y1 = -(2.0 / ORC_PI) * (j0 / x - lg * j1 - s_y1);
}

static void hankel_asym(double x, double nu, double *hre, double *him) {
  /* sum_k i^k a_k / x^k = P + iQ */
  double P = 1.0, Q = 0.0;
  double a = 1.0, prev = 1e300;
  const double mu = 4.0 * nu * nu;
  for (int k = 1; k < 200; ++k) {
    a *= (mu - (2.0 * k - 1) * (2.0 * k - 1)) / (8.0 * k * x);
    double t = fabs(a);
    if (t > prev) break;
    prev = t;
    switch (k & 3) {
      case 0: P += a; break;
      case 1: Q += a; break;
      case 2: P -= a; break;
      case 3: Q -= a; break;
    }
    if (t < 1e-18) break;
  }
  const double amp = sqrt(2.0 / (ORC_PI * x));
  /* e^{i(x - nu pi/2 - pi/4)}: combine sin/cos of x with the exact phase. */
  double s = sin(x), c = cos(x);
  double cr, ci; /* e^{-i(nu pi/2 + pi/4)} */
  const double r2 = 0.70710678118654752440;
  if (nu == 0.0) { cr = r2; ci = -r2; } else { cr = -r2; ci = -r2; }
  double er = c * cr - s * ci, ei = c * ci + s * cr;
  *hre = amp * (P * er - Q * ei);
  *him = amp * (P * ei + Q * er);
}

void orc_bessel01(double x, double *j0, double *j1, double *y0, double *y1) {
  if (x < 20.0) {
    bessel_miller(x, j0, j1, y0, y1);
  } else {
    double a, b;
    hankel_asym(x, 0.0, &a, &b); *j0 = a; *y0 = b;
    hankel_asym(x, 1.0, &a, &b); *j1 = a; *y1 = b;
  }
}

void orc_hankel01(double x, double *h0re, double *h0im, double *h1re, double *h1im) {
  double j0, j1, y0, y1;
  orc_bessel01(x, &j0, &j1, &y0, &y1);
  *h0re = j0; *h0im = y0; *h1re = j1; *h1im = y1;
}
