/* Precision template for the oracle FFT; included by orc_fft.c with
 * R (real type), SFX (name suffix) defined.  Self-sorting Stockham
 * formulation: stage with radix R over span Ns reads in[j + r*n/R] and writes
 * out[(j/Ns)*Ns*R + j%Ns + r*Ns] after the twiddle exp(sign 2 pi i r (j%Ns)/(Ns R)).
 * The reference calls cuFFT / MKL DFTI here (B/CUDA_GMRES/Fast_matvec.cuh:13-26,
 * B/MKL_GMRES_test_09122019/source/mkl_based_functions.h:5-15); both compute
 * the same unnormalised DFT, pinned against numpy.fft in tests/test_oracle.py. */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)
#define CX CAT(cx, SFX)

typedef struct { R re, im; } CX;

static int FN(factorize)(int n, int *f) {
  int k = 0;
  while (n % 4 == 0 && n > 4) { f[k++] = 4; n /= 4; }
  static const int primes[] = {2, 3, 5, 7, 11, 13};
  for (int p = 0; p < 6; ++p)
    while (n % primes[p] == 0) { f[k++] = primes[p]; n /= primes[p]; }
  for (int p = 17; n > 1; p += 2)
    while (n % p == 0) { f[k++] = p; n /= p; }
  return k;
}

/* One 1-D transform of length n, contiguous, out-of-place ping-pong. */
static void FN(fft1)(int n, CX *data, CX *work, int sign, const double *twr, const double *twi) {
  int f[64];
  const int nf = FN(factorize)(n, f);
  CX *in = data, *out = work;
  int Ns = 1;
  CX v[64], t[64];
  CX *vbig = NULL, *tbig = NULL;
  for (int s = 0; s < nf; ++s) {
    const int Rr = f[s];
    CX *vv = v, *tt = t;
    if (Rr > 64) {
      vbig = (CX *)malloc(sizeof(CX) * Rr);
      tbig = (CX *)malloc(sizeof(CX) * Rr);
      vv = vbig; tt = tbig;
    }
    const int nR = n / Rr;
    for (int j = 0; j < nR; ++j) {
      const int k = j % Ns;
      for (int r = 0; r < Rr; ++r) {
        CX a = in[j + r * nR];
        /* twiddle exp(sign*2*pi*i*r*k/(Ns*Rr)) = table entry r*k*(n/(Ns*Rr)) */
        const long q = (long)r * k * (n / (Ns * Rr));
        const double wr = twr[q], wi = -sign * twi[q];
        vv[r].re = (R)(a.re * wr - a.im * wi);
        vv[r].im = (R)(a.re * wi + a.im * wr);
      }
      if (Rr == 2) {
        tt[0].re = vv[0].re + vv[1].re; tt[0].im = vv[0].im + vv[1].im;
        tt[1].re = vv[0].re - vv[1].re; tt[1].im = vv[0].im - vv[1].im;
      } else if (Rr == 4) {
        /* X_q = sum_r v_r w^{rq}, w = exp(sign 2 pi i/4) = sign*i */
        const R ar = vv[0].re + vv[2].re, ai = vv[0].im + vv[2].im;
        const R br = vv[0].re - vv[2].re, bi = vv[0].im - vv[2].im;
        const R cr = vv[1].re + vv[3].re, ci = vv[1].im + vv[3].im;
        const R dr = vv[1].re - vv[3].re, di = vv[1].im - vv[3].im;
        /* (sign*i)*d */
        const R sdr = -sign * di, sdi = sign * dr;
        tt[0].re = ar + cr; tt[0].im = ai + ci;
        tt[2].re = ar - cr; tt[2].im = ai - ci;
        tt[1].re = br + sdr; tt[1].im = bi + sdi;
        tt[3].re = br - sdr; tt[3].im = bi - sdi;
      } else {
        for (int qq = 0; qq < Rr; ++qq) {
          double sr = 0, si = 0;
          for (int r = 0; r < Rr; ++r) {
            const long idx = ((long)r * qq % Rr) * (n / Rr);
            const double wr = twr[idx], wi = -sign * twi[idx];
            sr += vv[r].re * wr - vv[r].im * wi;
            si += vv[r].re * wi + vv[r].im * wr;
          }
          tt[qq].re = (R)sr;
          tt[qq].im = (R)si;
        }
      }
      const int base = (j / Ns) * Ns * Rr + k;
      for (int r = 0; r < Rr; ++r) out[base + r * Ns] = tt[r];
    }
    if (vbig) { free(vbig); free(tbig); vbig = tbig = NULL; }
    CX *tmp = in; in = out; out = tmp;
    Ns *= Rr;
  }
  if (in != data) memcpy(data, in, sizeof(CX) * n);
}

/* twiddle table exp(-2 pi i q / n), q in [0, n), FP64 */
static void FN(twiddles)(int n, double *twr, double *twi) {
  for (int q = 0; q < n; ++q) {
    const double a = -2.0 * 3.14159265358979323846 * (double)q / (double)n;
    twr[q] = cos(a);
    twi[q] = sin(a);
  }
}

void FN(orc_fft_)(int n, R *data, int sign) {
  double *twr = (double *)malloc(sizeof(double) * n * 2);
  double *twi = twr + n;
  CX *work = (CX *)malloc(sizeof(CX) * n);
  FN(twiddles)(n, twr, twi);
  FN(fft1)(n, (CX *)data, work, sign, twr, twi);
  free(work);
  free(twr);
}

void FN(orc_fft2_)(int M, R *data, int sign) {
  double *twr = (double *)malloc(sizeof(double) * M * 2);
  double *twi = twr + M;
  CX *work = (CX *)malloc(sizeof(CX) * M);
  CX *col = (CX *)malloc(sizeof(CX) * M);
  CX *d = (CX *)data;
  FN(twiddles)(M, twr, twi);
  for (int i = 0; i < M; ++i) FN(fft1)(M, d + (long)i * M, work, sign, twr, twi);
  for (int j = 0; j < M; ++j) {
    for (int i = 0; i < M; ++i) col[i] = d[(long)i * M + j];
    FN(fft1)(M, col, work, sign, twr, twi);
    for (int i = 0; i < M; ++i) d[(long)i * M + j] = col[i];
  }
  free(col);
  free(work);
  free(twr);
}

#undef FN
#undef CX
#undef CAT
#undef CAT2
