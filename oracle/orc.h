/*
 * CPU oracle for the Helmholtz GFIEM GMRES path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path (cugmres_amd/) never calls it.
 *
 * This is a plain-C restatement of the reference's algorithm, written from the
 * reference's behaviour (all citations relative to the reference checkout,
 * B/ = "Benchmarks of CUDA-based GMRES Solver ... Photonic Devices/"):
 *   - Green's block + circulant embedding   B/MKL_GMRES_test_09122019/source/loop_functions.h:128-168
 *   - kernel spectrum Ghat = FFT2(G)        B/CUDA_GMRES/launch_GMRES.cuh:138-142
 *   - FFT matvec x - G(chi x) and adjoint   B/CUDA_GMRES/Fast_matvec.cuh:1-56,
 *                                           B/CUDA_GMRES/GMRES_kernels.cuh:98-149,286-326
 *   - plane-wave RHS and r0 = b - A x0      B/CUDA_GMRES/GMRES_kernels.cuh:169-244,557-572
 *   - reference-form GMRES (MGS, non-unitary
 *     rotations, LU least squares)          B/CUDA_GMRES/GMRES.cuh:22-511
 *   - standard restarted GMRES(m) with CGS2, unitary Givens and back
 *     substitution, and FP64 iterative refinement: textbook algorithms (Saad,
 *     "Iterative Methods for Sparse Linear Systems", 2nd ed., Alg. 6.9/6.11;
 *     Giraud et al. 2005 for CGS2) that the volume-integral operator requires.
 *
 * Parity pins for every function are listed in oracle/README.md.
 */
#ifndef CGM_ORACLE_H
#define CGM_ORACLE_H
#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- special functions (orc_special.c) ---------------------------------- */
/* J0, J1, Y0, Y1 of real x > 0 in FP64. */
void orc_bessel01(double x, double *j0, double *j1, double *y0, double *y1);
/* Hankel functions of the first kind H0(x), H1(x), returned as re/im pairs. */
void orc_hankel01(double x, double *h0re, double *h0im, double *h1re, double *h1im);

/* ---- problem generators (orc_green.c) ----------------------------------- */
/* Reference-mode per-cell wave number 2*3.14f/(N/6.f), evaluated in float as
 * B/CUDA_GMRES/GMRES_kernels.cuh:12 does. */
float orc_ref_wave_number(int N);
/* Reference-mode Green's block, N*N row-major, float rounding as
 * B/MKL_GMRES_test_09122019/source/loop_functions.h:134-143. (re,im) pairs. */
void orc_ref_green_block_f32(int N, float *blk);
/* Reference-mode embedded array G_prep ((2N-1)^2, "re im" pairs), laid out as
 * loop_functions.h:150-163. */
void orc_ref_gprep_f32(int N, float *gprep);
/* Baseline-mode block of k0^2*G in FP64: off-diagonal kh^2 (i/4) H0(kh*rho),
 * self term (i pi ka/2) H1(ka) - 1 with ka = kh/sqrt(pi). (re,im) pairs. */
void orc_baseline_green_block(int N, double kh, double *blk);
/* Reference-form block in FP64 for any wave number and half-cell shifts:
 * (i/4) H0(k sqrt((i - sr)^2 + (j - sc)^2)); sr = 0.5 is the C++ orientation
 * (loop_functions.h:139), sc = 0.5 the Python one (B/PyGMRES/GMRES.py:23-27). */
void orc_shifted_green_block(int N, double k, double sr, double sc, double *blk);
/* Extract the Toeplitz generator from a (2N-1)^2 embedded array into an N*N
 * block; returns 0 when the array is even-symmetric in both offsets. */
int orc_gprep_to_block(int N, const float *gprep, double *blk);
/* Even circulant embedding of an N*N generator block into M*M (M >= 2N-1). */
void orc_embed_block(int N, int M, const double *blk, double *E);
/* Reference disk mask (row N/3, col N/2, radius N/6). */
void orc_ref_cylinder_mask(int N, unsigned char *mask);
/* Reference plane wave exp(-i*fl(k*i)) (alpha = 0) in float, (re,im). */
void orc_ref_plane_wave_f32(int N, float *b);
/* exp(-i*kh*(i cos th + j sin th)) in FP64, (re,im). */
void orc_plane_wave(int N, double kh, double theta, double *b);

/* ---- FFT (orc_fft.c) ----------------------------------------------------- */
/* In-place 1-D/2-D complex DFT, unnormalised, sign -1 forward, +1 inverse.
 * Any length (mixed radix 2/3/4/5 plus a direct stage for other primes). */
void orc_fft_c128(int n, double *data, int sign);
void orc_fft2_c128(int M, double *data, int sign);
void orc_fft_c64(int n, float *data, int sign);
void orc_fft2_c64(int M, float *data, int sign);

/* ---- operator + solvers (orc_gmres.c) ------------------------------------ */
typedef struct {
  int N, M;
  double *ghat;   /* M*M complex FP64 spectrum of the embedded kernel */
  float *ghat32;  /* same, rounded to complex64 */
  double *tw;     /* FP64 twiddle table exp(-2 pi i q/M), (re,im) pairs */
} orc_op;
/* Builds the spectrum of an N*N generator block embedded in M*M. */
int orc_op_init(orc_op *op, int N, int M, const double *blk);
void orc_op_free(orc_op *op);
/* mode 0: y = x - G(chi x);  1 (adjoint): y = x - chi G(x);  2: y = G(x) */
void orc_matvec_c128(const orc_op *op, const double *chi, const double *x, double *y, int mode);
void orc_matvec_c64(const orc_op *op, const float *chi, const float *x, float *y, int mode);
/* Brute-force dense BTTB product, O(N^4), for small N: same modes. */
void orc_matvec_dense(int N, const double *blk, const double *chi, const double *x, double *y, int mode);

typedef struct {
  int restart;          /* m */
  int max_cycles;       /* restart cap */
  double tol;           /* relative true residual */
  int adjoint;          /* 1: solve A^T x = b */
  int reference_form;   /* 1: reference rotations + LU, fixed m steps, no restart */
} orc_gmres_opts;

typedef struct {
  int iterations;       /* Arnoldi steps done in total */
  int cycles;
  int converged;
  double rel_residual;  /* true ||b - A x|| / ||b|| at exit (FP64) */
  double beta0;         /* initial residual norm */
} orc_gmres_info;

/* Restarted GMRES(m) with CGS2 + unitary Givens + back substitution, all FP64.
 * x holds x0 on entry.  res_hist (optional, length >= max steps + 1) receives
 * the Givens residual estimates |g_{i+1}| (absolute). */
int orc_gmres_c128(const orc_op *op, const double *chi, const double *b, double *x,
                   const orc_gmres_opts *o, orc_gmres_info *info, double *res_hist);
/* Same algorithm with complex64 vectors/FFTs and FP64 Givens (CPU baseline). */
int orc_gmres_c64(const orc_op *op, const float *chi, const float *b, float *x,
                  const orc_gmres_opts *o, orc_gmres_info *info, double *res_hist);
/* Reference-form GMRES, B/CUDA_GMRES/GMRES.cuh:22-446 (FP64 arithmetic):
 * r0 = b - A x0, MGS, rotation J = [[c, s], [-conj(s), c]] with c = h0/d,
 * s = h1/d, residual estimates beta*|Jtotal[i+1,0]|, LU solve of the m x m
 * system (Jtotal H)[0:m,0:m] y = beta Jtotal[0:m,0], x += V y.
 * Optional dumps: H (maxiter+1)*maxiter, Jtotal (maxiter+1)^2 row-major,
 * V (maxiter+1)*N^2, residual estimates (maxiter+1). */
int orc_gmres_reference(const orc_op *op, const double *chi, const double *b, double *x,
                        int maxiter, int adjoint, double *H_out, double *J_out,
                        double *V_out, double *res_out);
/* FP64 iterative refinement: outer c128 residual, inner complex64 GMRES(m). */
int orc_gmres_refine(const orc_op *op, const double *chi, const double *b, double *x,
                     int inner_restart, double inner_tol, int inner_max_cycles,
                     double outer_tol, int max_outer, orc_gmres_info *info);

/* Wall-clock helper for the CPU baseline. */
double orc_now(void);

#ifdef __cplusplus
}
#endif
#endif
