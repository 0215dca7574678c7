/*
 * Oracle operator and solvers (test-only).  See orc.h for the reference
 * functions each entry restates.
 */
#define _POSIX_C_SOURCE 199309L
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "orc.h"

#define R double
#define SFX c128
#include "orc_fft_impl.h"
#include "orc_solver_impl.h"
#undef R
#undef SFX

#define R float
#define SFX c64
#include "orc_fft_impl.h"
#include "orc_solver_impl.h"
#undef R
#undef SFX

double orc_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Ghat = FFT2 of the even embedding, computed in FP64
 * (B/CUDA_GMRES/launch_GMRES.cuh:138-142 does it in FP32 with cuFFT). */
int orc_op_init(orc_op *op, int N, int M, const double *blk) {
  if (M < 2 * N - 1) return 1;
  op->N = N;
  op->M = M;
  op->ghat = (double *)malloc(sizeof(double) * 2 * (size_t)M * M);
  op->ghat32 = (float *)malloc(sizeof(float) * 2 * (size_t)M * M);
  op->tw = (double *)malloc(sizeof(double) * 2 * (size_t)M);
  for (int q = 0; q < M; ++q) {
    const double a = -2.0 * 3.14159265358979323846 * (double)q / (double)M;
    op->tw[q] = cos(a);
    op->tw[M + q] = sin(a);
  }
  orc_embed_block(N, M, blk, op->ghat);
  orc_fft2_c128(M, op->ghat, -1);
  for (long q = 0; q < 2L * M * M; ++q) op->ghat32[q] = (float)op->ghat[q];
  return 0;
}

void orc_op_free(orc_op *op) {
  free(op->ghat);
  free(op->ghat32);
  free(op->tw);
  op->ghat = NULL;
  op->ghat32 = NULL;
  op->tw = NULL;
}

/* Direct BTTB product: (G v)[p] = sum_q g(|i_p - i_q|, |j_p - j_q|) v[q]. */
void orc_matvec_dense(int N, const double *blk, const double *chi, const double *x_,
                      double *y_, int mode) {
  const long n = (long)N * N;
  const double complex *x = (const double complex *)x_;
  double complex *y = (double complex *)y_;
  const double complex *g = (const double complex *)blk;
  double complex *u = (double complex *)malloc(sizeof(double complex) * n);
  for (long q = 0; q < n; ++q) u[q] = (mode == 0) ? chi[q] * x[q] : x[q];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      double complex s = 0;
      for (int a = 0; a < N; ++a)
        for (int c = 0; c < N; ++c) {
          const int di = abs(i - a), dj = abs(j - c);
          s += g[(long)di * N + dj] * u[(long)a * N + c];
        }
      const long p = (long)i * N + j;
      if (mode == 2) y[p] = s;
      else if (mode == 1) y[p] = x[p] - chi[p] * s;
      else y[p] = x[p] - s;
    }
  free(u);
}

/* Reference-form GMRES, B/CUDA_GMRES/GMRES.cuh:22-446 and
 * B/MKL_GMRES_test_09122019/source/MKL_GMRES.h:3-395, in FP64.
 * H is (maxiter+1) x maxiter row-major (ld maxiter) as dev_H_; Jtotal is
 * (maxiter+1)^2 row-major (ld maxiter+1).  The LU solve uses partial
 * pivoting as getrf/getrs do (GMRES.cuh:450-511). */
int orc_gmres_reference(const orc_op *op, const double *chi, const double *b_, double *x_,
                        int maxiter, int adjoint, double *H_out, double *J_out,
                        double *V_out, double *res_out) {
  const long n = (long)op->N * op->N;
  const int m = maxiter, m1 = maxiter + 1;
  const double complex *b = (const double complex *)b_;
  double complex *x = (double complex *)x_;
  double complex *V = (double complex *)malloc(sizeof(double complex) * n * m1);
  double complex *H = (double complex *)calloc((size_t)m1 * m, sizeof(double complex));
  double complex *J = (double complex *)calloc((size_t)m1 * m1, sizeof(double complex));
  double complex *Jn = (double complex *)calloc((size_t)m1 * m1, sizeof(double complex));
  double *res = (double *)calloc(m1, sizeof(double));
  /* r0 = b - A x0 (GMRES_kernels.cuh:179-208) */
  orc_matvec_c128(op, chi, x_, (double *)V, adjoint ? 1 : 0);
  for (long e = 0; e < n; ++e) V[e] = b[e] - V[e];
  double beta = 0;
  for (long e = 0; e < n; ++e) beta += creal(V[e] * conj(V[e]));
  beta = sqrt(beta);
  res[0] = beta;
  for (long e = 0; e < n; ++e) V[e] /= beta;
  for (int i = 0; i < m1; ++i) J[(long)i * m1 + i] = 1.0;
  int GMRES_i = 0;
  for (int i = 0; i < m; ++i) {
    double complex *vi = V + (long)i * n, *vn = V + (long)(i + 1) * n;
    orc_matvec_c128(op, chi, (const double *)vi, (double *)vn, adjoint ? 1 : 0);
    /* modified Gram-Schmidt (GMRES.cuh:289-303) */
    for (int j = 0; j <= i; ++j) {
      double complex h = 0;
      const double complex *vj = V + (long)j * n;
      for (long e = 0; e < n; ++e) h += conj(vj[e]) * vn[e];
      H[(long)j * m + i] = h;
      for (long e = 0; e < n; ++e) vn[e] -= h * vj[e];
    }
    double hn = 0;
    for (long e = 0; e < n; ++e) hn += creal(vn[e] * conj(vn[e]));
    hn = sqrt(hn);
    H[(long)(i + 1) * m + i] = hn;
    for (long e = 0; e < n; ++e) vn[e] /= hn;
    /* Htemp = Jtotal H (column i is all that the rotation reads) */
    double complex h0 = 0, h1 = 0;
    for (int r = 0; r <= i + 1; ++r) {
      h0 += J[(long)i * m1 + r] * H[(long)r * m + i];
      h1 += J[(long)(i + 1) * m1 + r] * H[(long)r * m + i];
    }
    /* set_4_Givens_rotation_matrix_elements_kernel (GMRES_kernels.cuh:453-509) */
    const double d = sqrt(creal(h0 * conj(h0)) + creal(h1 * conj(h1)));
    const double complex c = h0 / d, s = h1 / d;
    /* Jtotal = J Jtotal acting on rows i, i+1 */
    memcpy(Jn, J, sizeof(double complex) * m1 * m1);
    for (int col = 0; col < m1; ++col) {
      const double complex a = J[(long)i * m1 + col], bb = J[(long)(i + 1) * m1 + col];
      Jn[(long)i * m1 + col] = c * a + s * bb;
      Jn[(long)(i + 1) * m1 + col] = -conj(s) * a + c * bb;
    }
    memcpy(J, Jn, sizeof(double complex) * m1 * m1);
    res[i + 1] = beta * cabs(J[(long)(i + 1) * m1]);
    GMRES_i = i + 1;
  }
  /* HH = (Jtotal H)[0:m, 0:m], cc = beta Jtotal[0:m, 0] (GMRES.cuh:377-397) */
  const int k = GMRES_i;
  double complex *A = (double complex *)calloc((size_t)k * k, sizeof(double complex));
  double complex *cc = (double complex *)calloc(k, sizeof(double complex));
  for (int r = 0; r < k; ++r) {
    for (int col = 0; col < k; ++col) {
      double complex s = 0;
      for (int t = 0; t <= k; ++t) s += J[(long)r * m1 + t] * H[(long)t * m + col];
      A[(long)r * k + col] = s;
    }
    cc[r] = beta * J[(long)r * m1];
  }
  /* LU with partial pivoting, then forward/back substitution */
  for (int col = 0; col < k; ++col) {
    int p = col;
    for (int r = col + 1; r < k; ++r)
      if (cabs(A[(long)r * k + col]) > cabs(A[(long)p * k + col])) p = r;
    if (p != col) {
      for (int t = 0; t < k; ++t) {
        double complex tmp = A[(long)col * k + t];
        A[(long)col * k + t] = A[(long)p * k + t];
        A[(long)p * k + t] = tmp;
      }
      double complex tmp = cc[col]; cc[col] = cc[p]; cc[p] = tmp;
    }
    for (int r = col + 1; r < k; ++r) {
      const double complex f = A[(long)r * k + col] / A[(long)col * k + col];
      for (int t = col; t < k; ++t) A[(long)r * k + t] -= f * A[(long)col * k + t];
      cc[r] -= f * cc[col];
    }
  }
  for (int r = k - 1; r >= 0; --r) {
    double complex s = cc[r];
    for (int t = r + 1; t < k; ++t) s -= A[(long)r * k + t] * cc[t];
    cc[r] = s / A[(long)r * k + r];
  }
  for (int j = 0; j < k; ++j)
    for (long e = 0; e < n; ++e) x[e] += cc[j] * V[(long)j * n + e];
  if (H_out) memcpy(H_out, H, sizeof(double complex) * m1 * m);
  if (J_out) memcpy(J_out, J, sizeof(double complex) * m1 * m1);
  if (V_out) memcpy(V_out, V, sizeof(double complex) * n * m1);
  if (res_out) memcpy(res_out, res, sizeof(double) * m1);
  free(cc); free(A); free(res); free(Jn); free(J); free(H); free(V);
  return 0;
}

/* Outer FP64 residual, inner complex64 GMRES on the normalised residual. */
int orc_gmres_refine(const orc_op *op, const double *chi, const double *b_, double *x_,
                     int inner_restart, double inner_tol, int inner_max_cycles,
                     double outer_tol, int max_outer, orc_gmres_info *info) {
  const long n = (long)op->N * op->N;
  const double complex *b = (const double complex *)b_;
  double complex *x = (double complex *)x_;
  double complex *r = (double complex *)malloc(sizeof(double complex) * n);
  float *chi32 = (float *)malloc(sizeof(float) * n);
  float *r32 = (float *)malloc(sizeof(float) * 2 * n);
  float *d32 = (float *)malloc(sizeof(float) * 2 * n);
  for (long e = 0; e < n; ++e) chi32[e] = (float)chi[e];
  double bn = 0;
  for (long e = 0; e < n; ++e) bn += creal(b[e] * conj(b[e]));
  bn = sqrt(bn);
  orc_gmres_opts o = {inner_restart, inner_max_cycles, inner_tol, 0, 0};
  info->iterations = 0;
  info->cycles = 0;
  info->converged = 0;
  for (int it = 0; it <= max_outer; ++it) {
    orc_matvec_c128(op, chi, x_, (double *)r, 0);
    double rn = 0;
    for (long e = 0; e < n; ++e) { r[e] = b[e] - r[e]; rn += creal(r[e] * conj(r[e])); }
    rn = sqrt(rn);
    info->rel_residual = bn > 0 ? rn / bn : 0.0;
    if (it == 0) info->beta0 = rn;
    if (rn <= outer_tol * bn) { info->converged = 1; break; }
    if (it == max_outer) break;
    for (long e = 0; e < n; ++e) {
      r32[2 * e] = (float)(creal(r[e]) / rn);
      r32[2 * e + 1] = (float)(cimag(r[e]) / rn);
      d32[2 * e] = d32[2 * e + 1] = 0.0f;
    }
    orc_gmres_info ii;
    orc_gmres_c64(op, chi32, r32, d32, &o, &ii, NULL);
    info->iterations += ii.iterations;
    info->cycles += 1;
    for (long e = 0; e < n; ++e) x[e] += rn * (d32[2 * e] + I * d32[2 * e + 1]);
  }
  free(d32); free(r32); free(chi32); free(r);
  return info->converged ? 0 : 4;
}
