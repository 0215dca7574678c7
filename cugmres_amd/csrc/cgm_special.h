// FP64 Bessel/Hankel functions of orders 0 and 1, host and device.
//
// The reference evaluates the Green's function on the host with libm j0/y0
// (B/MKL_GMRES_test_09122019/source/loop_functions.h:139-142); the spectrum
// here is generated on the GPU in FP64, so the functions are written from
// published mathematics:
//   x < 20 : Miller backward recurrence for J_n normalised by
//            J0 + 2 sum_k J_2k = 1 (A&S 9.1.87), Neumann series for Y0
//            (A&S 9.1.88) and its term-by-term derivative for Y1 = -Y0'.
//   x >= 20: Hankel asymptotic expansion (DLMF 10.17.5) summed to its
//            smallest term.
#pragma once
#include <math.h>

namespace cgm {

#define CGM_PI 3.14159265358979323846
#define CGM_EULER 0.57721566490153286061

__host__ __device__ inline void bessel01_small(double x, double& J0, double& J1, double& Y0,
                                               double& Y1) {
  // Backward recurrence without storing all orders: keep running sums.
  int n0 = (int)(x + 12.0 * cbrt(x) + 40.0);
  if (n0 & 1) n0++;
  double jp1 = 0.0;    // J_{n+1}
  double jn = 1e-30;   // J_n
  double s_norm = 0.0, s_y0 = 0.0, s_y1 = 0.0;
  // For even n = 2k the Y1 term needs J_{2k-1} - J_{2k+1}; J_{2k+1} is jp1
  // when jn = J_{2k}, and J_{2k-1} is produced by the next recurrence step.
  double pend = 0.0;
  int pend_k = 0;
  double j1 = 0.0;
  for (int n = n0; n >= 1; --n) {
    if ((n & 1) == 0) {
      const int k = n >> 1;
      const double sg = (k & 1) ? -1.0 : 1.0;
      s_norm += jn;
      s_y0 += sg * jn / k;
      pend = jp1;
      pend_k = k;
    }
    const double jm = (2.0 * n / x) * jn - jp1;  // J_{n-1}
    if ((n & 1) == 0) {
      const double sg = (pend_k & 1) ? -1.0 : 1.0;
      s_y1 += sg * (jm - pend) / pend_k;
    }
    jp1 = jn;
    jn = jm;
    if (n == 2) j1 = jn;  // jn = J_1 after the n = 2 step
    if (fabs(jn) > 1e250) {
      jn *= 1e-250; jp1 *= 1e-250; s_norm *= 1e-250; s_y0 *= 1e-250; s_y1 *= 1e-250;
      pend *= 1e-250; j1 *= 1e-250;
    }
  }
  const double norm = jn + 2.0 * s_norm;
  J0 = jn / norm;
  J1 = j1 / norm;
  s_y0 /= norm;
  s_y1 /= norm;
  const double lg = log(0.5 * x) + CGM_EULER;
  Y0 = (2.0 / CGM_PI) * (lg * J0 - 2.0 * s_y0);
  Y1 = -(2.0 / CGM_PI) * (J0 / x - lg * J1 - s_y1);
}

__host__ __device__ inline void hankel_asym(double x, int nu, double& hre, double& him) {
  double P = 1.0, Q = 0.0, a = 1.0, prev = 1e300;
  const double mu = 4.0 * nu * nu;
  for (int k = 1; k < 200; ++k) {
    a *= (mu - (2.0 * k - 1) * (2.0 * k - 1)) / (8.0 * k * x);
    const double t = fabs(a);
    if (t > prev) break;
    prev = t;
    switch (k & 3) {
      case 0: P += a; break;
      case 1: Q += a; break;
      case 2: P -= a; break;
      default: Q -= a; break;
    }
    if (t < 1e-18) break;
  }
  const double amp = sqrt(2.0 / (CGM_PI * x));
  double s, c;
  sincos(x, &s, &c);
  const double r2 = 0.70710678118654752440;
  const double cr = nu == 0 ? r2 : -r2, ci = -r2;
  const double er = c * cr - s * ci, ei = c * ci + s * cr;
  hre = amp * (P * er - Q * ei);
  him = amp * (P * ei + Q * er);
}

// J0, J1, Y0, Y1 at x > 0.
__host__ __device__ inline void bessel01(double x, double& J0, double& J1, double& Y0, double& Y1) {
  if (x < 20.0) {
    bessel01_small(x, J0, J1, Y0, Y1);
  } else {
    hankel_asym(x, 0, J0, Y0);
    hankel_asym(x, 1, J1, Y1);
  }
}

}  // namespace cgm
