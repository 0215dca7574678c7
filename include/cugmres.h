/*
 * cugmres.h -- C ABI of the gfx950 GMRES solver for the 2-D Helmholtz
 * Green's-function integral equation
 *
 *     A x = x - G (chi .* x),      G block-Toeplitz with Toeplitz blocks,
 *
 * applied through a zero-padded M x M circulant embedding (M >= 2N-1).
 *
 * Every entry point takes plain pointers and sizes.  Vectors are N*N
 * interleaved complex values, row-major, index i*N + j with i the row (the
 * direction of the reference's incident wave).  Device pointers are HIP
 * device memory on the context's device.  Functions return a cgm_status;
 * none of them exits the process.
 *
 * Reference interfaces replaced (paths relative to the reference checkout,
 * B/ = "Benchmarks of CUDA-based GMRES Solver for Toeplitz and Hankel
 * Matrices and Applications to Topology Optimization of Photonic Devices/"):
 *   cufftHandle + cublasHandle_t + cusolverDnHandle_t  -> cgm_ctx_t
 *        (B/CUDA_GMRES/launch_GMRES.cuh:93-101)
 *   cufftExecC2C(plan, G_prep, FORWARD)                 -> cgm_spectrum_from_gprep
 *        (B/CUDA_GMRES/launch_GMRES.cuh:138-142)
 *   Green_matrix_create_MKL + get_gamma_array           -> cgm_spectrum_reference
 *        (B/MKL_GMRES_test_09122019/source/loop_functions.h:128-168,
 *         B/MKL_GMRES_test_09122019/source/mkl_based_functions.h:19-34)
 *   G_x_fft_matvec (+ _2D_to_1D_kernel)                 -> cgm_matvec, cgm_G_x_fft_matvec
 *        (B/CUDA_GMRES/Fast_matvec.cuh:1-56, B/CUDA_GMRES/GMRES_kernels.cuh:286-326)
 *   Fast_GMRES_with_CUDA                                -> cgm_gmres, cgm_Fast_GMRES_with_CUDA
 *        (B/CUDA_GMRES/GMRES.cuh:22-446)
 *   cuSolve_LES                                         -> cgm_cuSolve_LES
 *        (B/CUDA_GMRES/GMRES.cuh:450-511)
 *   init_x0_kernel                                      -> cgm_reference_plane_wave
 *        (B/CUDA_GMRES/GMRES_kernels.cuh:557-572)
 */
#ifndef CUGMRES_H
#define CUGMRES_H

#include <stddef.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGM_VERSION_MAJOR 0
#define CGM_VERSION_MINOR 1

/* Layout-identical to cuComplex/hipFloatComplex ({float x, y}) and to the
 * reference's complex8 (B/MKL_GMRES_test_09122019/source/parameters.h:50). */
typedef struct { float re, im; } cgm_complex64;
typedef struct { double re, im; } cgm_complex128;

typedef struct cgm_context* cgm_ctx_t;

typedef enum {
  CGM_SUCCESS = 0,
  CGM_ERROR_INVALID_VALUE = 1,
  CGM_ERROR_DEVICE = 2,
  CGM_ERROR_ALLOC = 3,
  CGM_ERROR_NOT_CONVERGED = 4,
  CGM_ERROR_NONFINITE = 5,
  CGM_ERROR_UNSUPPORTED = 7,
  CGM_ERROR_NOT_SYMMETRIC = 8
} cgm_status;

/* Operator modes (cgm_matvec). */
typedef enum {
  CGM_OP_FORWARD = 0,  /* y = x - G(chi x)          (B/CUDA_GMRES/Fast_matvec.cuh:1-28)  */
  CGM_OP_ADJOINT = 1,  /* y = x - chi G(x) = A^T x  (B/CUDA_GMRES/Fast_matvec.cuh:31-56) */
  CGM_OP_GREEN = 2     /* y = G(x)                  (B/PyGMRES/GMRES.py:30-44)          */
} cgm_op_mode;

const char* cgm_status_string(int status);
/* Message of the last error raised in this thread ("" if none). */
const char* cgm_last_error(void);
void cgm_version(int* major, int* minor);

/* ---- context ------------------------------------------------------------ */
/* One context per (grid size N, device, host thread).  Chooses the embedding
 * size M (smallest of 2^a, 3*2^a >= 2N-1; N <= 2048) and owns the FFT
 * twiddle tables and the scratch of one matvec.  Replaces the reference's
 * cufftPlan2d/cublasCreate/cusolverDnCreate (B/CUDA_GMRES/launch_GMRES.cuh:95-101). */
int cgm_create(cgm_ctx_t* ctx, int N, int device);
int cgm_destroy(cgm_ctx_t ctx);
/* All work of the context is enqueued on this HIP stream (hipStream_t passed
 * as void*; NULL = the default stream). */
int cgm_set_stream(cgm_ctx_t ctx, void* stream);
int cgm_get_embedding(cgm_ctx_t ctx, int* M);
/* Autotuner (run by cgm_create): times the work-group tiles of the three FFT
 * passes (rows per work-group of passes 1 and 3, columns per work-group of
 * pass 2, each in {1, 2, 4, 8} within 1024 threads and 64 KiB of LDS) with
 * HIP events and keeps the fastest.  Call again after cgm_set_stream to tune
 * on another stream. */
int cgm_autotune(cgm_ctx_t ctx);
/* tiles[0..2] = rows of pass 1, columns of pass 2, rows of pass 3; us[0..2]
 * (optional) = their tuned times in microseconds. */
int cgm_get_tiles(cgm_ctx_t ctx, int* tiles, float* us);
/* The autotuner's radix plan at M = 1024 (the N = 512 embedding): for each
 * pass q = 0, 1, 2 (passes 1-3), launches of at least wide_from_batch[q]
 * vectors use the register-resident one-wave kernel (one 64-lane wave per
 * 1024-point transform, radices 16, 16, 4, two LDS exchanges) instead of the
 * LDS-staged radix-8 stages (0: never; other M always use the radix-8
 * stages).  Passes 1 and 3 take it for the plain modes only (no residual,
 * no input scaling).  wide_us / lds_us (optional, 9 each, [pass][batch]) =
 * the two kernels' times at batches of 1, 4 and 16 vectors in microseconds. */
int cgm_get_radix_plan(cgm_ctx_t ctx, int* wide_from_batch, float* wide_us, float* lds_us);
/* Overrides that choice per pass (0: the radix-8 kernel; 1: the one-wave
 * kernel for every launch it supports).  Ignored at other M. */
int cgm_set_radix_plan(cgm_ctx_t ctx, const int* wide_from_batch);
/* Arnoldi kernels of the solvers: 0 (default) = fused steps with lagged CGS2
 * (one global reduction and two basis passes per step; N = 2^k in [64, 2048]
 * or 3 * 2^k in [96, 768], restart <= 64, other cases fall back to 1), with
 * 256-thread work-groups once a launch has at least 512 work-groups of 1024 / N
 * rows (batch * N * N >= 2^19) and 1,024-thread ones below (the two shapes
 * sum in different orders, so a
 * batched system and its single solve agree to rounding, bitwise under path
 * 2 or 3);
 * 1 = separate kernels with two-pass CGS2 (three basis passes per step);
 * 2 / 3 = fused steps with 256- / 1,024-thread work-groups always (the
 * 256-thread shape exists for N in {64, 96, 128, 192, 256, 512}). */
int cgm_set_arnoldi_path(cgm_ctx_t ctx, int path);
/* Number of complex entries of a kernel spectrum: (M/2+1)^2.  The spectrum of
 * an even circulant embedding is even in both frequencies, so only this
 * quarter is stored. */
int cgm_spectrum_size(cgm_ctx_t ctx, size_t* n_complex);

/* ---- kernel spectrum (FP64 on the device, stored complex64/complex128) ---- */
/* From the reference's embedded spatial array G_prep ((2N-1)^2 complex64,
 * device memory, layout of B/MKL_GMRES_test_09122019/source/loop_functions.h:150-163).
 * Returns CGM_ERROR_NOT_SYMMETRIC if G_prep is not even in both offsets. */
int cgm_spectrum_from_gprep(cgm_ctx_t ctx, const cgm_complex64* dev_gprep,
                            cgm_complex64* dev_gamma);
/* From an N x N generator block g[|di|, |dj|] (complex128, device memory). */
int cgm_spectrum_from_block(cgm_ctx_t ctx, const cgm_complex128* dev_block,
                            cgm_complex64* dev_gamma, cgm_complex128* dev_gamma64);
/* The reference's operator generated on the device in FP64:
 * g = (i/4) H0(k sqrt((i - row_shift)^2 + (j - col_shift)^2)); row_shift = 0.5,
 * col_shift = 0 is the C++ orientation, the reverse the Python one. */
int cgm_spectrum_reference(cgm_ctx_t ctx, double wave_number, double row_shift, double col_shift,
                           cgm_complex64* dev_gamma, cgm_complex128* dev_gamma64);
/* The benchmark operator k0^2 G for cells of size h, kh = k0 h:
 * off-diagonal kh^2 (i/4) H0(kh |r_p - r_q|/h), self term (i pi ka/2) H1(ka) - 1,
 * ka = kh / sqrt(pi).  Either output may be NULL. */
int cgm_spectrum_helmholtz(cgm_ctx_t ctx, double kh, cgm_complex64* dev_gamma,
                           cgm_complex128* dev_gamma64);

/* ---- operator ------------------------------------------------------------ */
/* y = A x (mode FORWARD), A^T x (ADJOINT) or G x (GREEN); chi is the real
 * contrast per cell (N*N float).  Asynchronous on the context stream. */
int cgm_matvec(cgm_ctx_t ctx, const cgm_complex64* dev_gamma, const float* dev_chi,
               const cgm_complex64* dev_x, cgm_complex64* dev_y, int mode);
/* The same for `batch` vectors of the context's N in one launch per FFT
 * pass: x and y are [batch][N*N], chi is [batch][N*N] (chi_stride = N*N) or
 * one map (chi_stride = 0).  No counterpart in the reference, whose matvec
 * takes one vector (B/CUDA_GMRES/Fast_matvec.cuh:1-28). */
int cgm_matvec_batched(cgm_ctx_t ctx, const cgm_complex64* dev_gamma, const float* dev_chi, long chi_stride,
                       const cgm_complex64* dev_x, cgm_complex64* dev_y, int batch, int mode);
/* cgm_matvec_batched without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * FFT intermediates live in a fixed device array of
 * cgm_noalloc_scratch_bytes() (64 MiB: 16 vectors at N = 512, 4 at
 * N = 1024), one per GPU, whose address cgm_create takes; a larger batch
 * fails with CGM_ERROR_INVALID_VALUE and launches nothing.  The caller still
 * provides x and y.  Calls on one GPU share its array: calls that overlap
 * there (several streams or host threads) race.  Same outputs, bit for bit,
 * as cgm_matvec_batched. */
int cgm_matvec_batched_noalloc(cgm_ctx_t ctx, const cgm_complex64* dev_gamma, const float* dev_chi,
                               long chi_stride, const cgm_complex64* dev_x, cgm_complex64* dev_y, int batch,
                               int mode);
size_t cgm_noalloc_scratch_bytes(void);
/* FP64 variant (spectrum from dev_gamma64, chi as double). */
int cgm_matvec_c128(cgm_ctx_t ctx, const cgm_complex128* dev_gamma64, const double* dev_chi,
                    const cgm_complex128* dev_x, cgm_complex128* dev_y, int mode);

/* ---- solver --------------------------------------------------------------- */
typedef struct {
  int restart;        /* Krylov dimension per cycle m (1..512)                    */
  int max_cycles;     /* restart cap (>= 1)                                      */
  double tol;         /* stop when the true ||b - A x|| <= tol ||b||             */
  int adjoint;        /* 1: solve A^T x = b                                      */
  int x0_zero;        /* 1: ignore dev_x on entry and start from 0               */
} cgm_gmres_params;

typedef struct {
  int iterations;          /* Arnoldi steps in total                            */
  int cycles;              /* restart cycles run                                */
  int converged;           /* 1 when the true relative residual <= tol          */
  double rel_residual;     /* true ||b - A x|| / ||b|| at exit (complex64 math) */
  double beta0;            /* ||b - A x0||                                      */
  double bnorm;            /* ||b||                                             */
} cgm_gmres_result;

/* Restarted GMRES(m): CGS2 orthogonalisation, unitary Givens rotations and
 * back substitution in FP64, everything device resident; one host
 * synchronisation per restart cycle.  dev_hist (optional, device, length
 * restart*max_cycles + 1) receives the absolute Givens residual estimates.
 * Synchronous on return.  Returns CGM_SUCCESS or CGM_ERROR_NOT_CONVERGED
 * (x then holds the last iterate). */
int cgm_gmres(cgm_ctx_t ctx, const cgm_complex64* dev_gamma, const float* dev_chi,
              const cgm_complex64* dev_b, cgm_complex64* dev_x, const cgm_gmres_params* params,
              cgm_gmres_result* result, double* dev_hist);

/* Batched GMRES(m): `batch` systems of the context's N in lockstep, every
 * kernel of the solve launched once for all of them (system = grid y index).
 * They share the kernel spectrum; chi is [batch][N*N] with chi_stride = N*N
 * (own maps) or one map with chi_stride = 0 (multiple right-hand sides of
 * one instance).  b and x are [batch][N*N]; results has batch entries.  A
 * converged system stops at its own tolerance, the others go on.  dev_hist
 * (optional) is [batch][restart*max_cycles + 1]: [0] = ||r0||, entry
 * c*restart + j = estimate after j steps of cycle c.  Power-of-two N in
 * [64, 1024] with restart <= 64 run batched; other sizes solve the systems
 * one after another.  Without dev_hist the launches of one restart cycle
 * (and the next cycle's residual) are captured once per call into a HIP
 * graph on the context's own non-blocking stream and replayed every cycle;
 * cgm_set_batched_graphs(ctx, 0) launches them one by one instead.  No
 * counterpart in the reference, which solves one system per call
 * (B/CUDA_GMRES/GMRES.cuh:22-35). */
int cgm_set_batched_graphs(cgm_ctx_t ctx, int enable);
int cgm_gmres_batched(cgm_ctx_t ctx, const cgm_complex64* dev_gamma, const float* dev_chi, long chi_stride,
                      const cgm_complex64* dev_b, cgm_complex64* dev_x, int batch,
                      const cgm_gmres_params* params, cgm_gmres_result* results, double* dev_hist);

/* FP64 iterative refinement: the true residual r = b - A x is formed with the
 * complex128 operator, the correction A d = r is solved by complex64
 * GMRES(inner_restart) to inner_tol, x += d, until ||r|| <= outer_tol ||b||. */
typedef struct {
  int inner_restart;
  int inner_max_cycles;
  double inner_tol;
  double outer_tol;
  int max_outer;
} cgm_refine_params;

typedef struct {
  int outer;               /* refinement steps                                  */
  int inner_iterations;    /* Arnoldi steps over all inner solves               */
  int converged;
  double rel_residual;     /* FP64 true relative residual at exit               */
} cgm_refine_result;

int cgm_gmres_refine(cgm_ctx_t ctx, const cgm_complex64* dev_gamma,
                     const cgm_complex128* dev_gamma64, const double* dev_chi,
                     const cgm_complex128* dev_b, cgm_complex128* dev_x,
                     const cgm_refine_params* params, cgm_refine_result* result);

/* ---- learned right preconditioner and flexible GMRES ----------------------- */
/* No counterpart in the reference (B/CUDA_GMRES/GMRES.cuh:22-446 runs plain
 * GMRES without a preconditioner).  A Fourier neural operator P with 4
 * spectral layers, 32 channels and 16 x 16 retained modes (kx in
 * {0..15, N-16..N-1}, ky in {0..15}) maps a unit-norm vector v and the
 * contrast map chi to P(v) = s_id v + FNO(chi * chi_scale, N v) / N.  BF16
 * activations and MFMA GEMMs, FP32 accumulation and spectral mixing.
 * Parameters are one FP32 device array of cgm_fno_num_params() values:
 *   W_in [3][32], b_in [32],
 *   per layer l = 0..3: W_l [32 in][32 out], b_l [32],
 *                       Wspec_l [32 kx][16 ky][32 in][32 out] complex (re, im),
 *   W_out [32][2], b_out [2], s_id (learned weight of the identity path). */
typedef struct cgm_fno* cgm_fno_t;

int cgm_fno_num_params(size_t* count);
/* Workspace for up to `batch` vectors at the context's N (64 <= N <= 1024,
 * N % 64 == 0). */
int cgm_fno_create(cgm_ctx_t ctx, int batch, float chi_scale, cgm_fno_t* net);
int cgm_fno_destroy(cgm_fno_t net);
/* Binds the parameter array (device, 16-byte aligned, kept by reference: it
 * must outlive the calls that use it) and converts the GEMM weights to BF16.
 * Call again after the parameters change. */
int cgm_fno_set_params(cgm_fno_t net, const float* dev_params);
/* z[b] = P(v[b]) for b < batch; v is [batch][N*N] unit-norm vectors, chi is
 * [batch][N*N].  train = 1 keeps the activations for cgm_fno_backward (chi
 * and v must then stay unchanged until it runs).  v and z must not alias. */
int cgm_fno_forward(cgm_fno_t net, const float* dev_chi, const cgm_complex64* dev_v,
                    cgm_complex64* dev_z, int batch, int train);
/* Gradient of a real loss L with respect to every parameter, given
 * dev_dz[b][p] = dL/dRe z + i dL/dIm z of the last training forward.
 * dev_grad has the parameter layout and is overwritten. */
int cgm_fno_backward(cgm_fno_t net, const cgm_complex64* dev_dz, float* dev_grad);

typedef struct {
  int restart;        /* Krylov dimension per cycle m (1..16384; above 64 the
                         long-recurrence path keeps the whole basis in HBM)      */
  int max_cycles;     /* restart cap                                              */
  double tol;         /* stop when the true ||b - A x|| <= tol ||b||              */
  int x0_zero;        /* 1: start from x = 0                                      */
  double stall_ratio; /* a preconditioned cycle that leaves more than this
                         fraction of the residual switches the rest of the
                         solve to plain GMRES (<= 0: never)                       */
} cgm_fgmres_params;

typedef struct {
  int iterations;        /* Arnoldi steps in total (both phases)                 */
  int cycles;            /* restart cycles in total                              */
  int converged;
  double rel_residual;   /* true relative residual at exit                       */
  double beta0;
  double bnorm;
  int precond_cycles;    /* cycles run with the preconditioner                   */
  int fell_back;         /* 1 when the stall rule switched to plain GMRES        */
} cgm_fgmres_result;

/* Right-preconditioned flexible GMRES(m) (Saad, SIAM J. Sci. Comput. 14
 * (1993) 461-469): z_j = P(v_j / ||v_j||), w = A z_j, orthogonalised against
 * V, x += Z y.  For power-of-two N in [64, 2048] and m <= 64 the network takes
 * the lagged once-orthogonalised vector and each step has one reduction
 * (fused lagged CGS2); otherwise two-pass CGS2.  Plain GMRES (cgm_gmres's
 * kernels) is the fallback and the referee: every cycle starts from the true
 * residual b - A x.  net = NULL applies the identity (then the iterates equal
 * cgm_gmres's).  dev_hist as in cgm_gmres.  No counterpart in the reference,
 * which has no preconditioner (B/CUDA_GMRES/GMRES.cuh:22-446). */
int cgm_fgmres(cgm_ctx_t ctx, const cgm_complex64* dev_gamma, const float* dev_chi,
               const cgm_complex64* dev_b, cgm_complex64* dev_x, cgm_fno_t net,
               const cgm_fgmres_params* params, cgm_fgmres_result* result, double* dev_hist);
/* `batch` systems of the context's N in lockstep: b and x are [batch][N*N],
 * chi is [batch][N*N] (chi_stride = N*N) or one map (chi_stride = 0); each
 * step applies the network to all live systems in one forward pass (net
 * created with at least `batch` vectors, or NULL for the identity).  Needs
 * the fused kernels (power-of-two N in [64, 2048], restart <= 64) and
 * stall_ratio = 0 (no fallback).  results[batch]; dev_hist, when given,
 * [batch][restart * max_cycles + 1].  Each system's iterates equal its own
 * cgm_fgmres solve.  No counterpart in the reference. */
int cgm_fgmres_batched(cgm_ctx_t ctx, const cgm_complex64* dev_gamma, const float* dev_chi, long chi_stride,
                       const cgm_complex64* dev_b, cgm_complex64* dev_x, cgm_fno_t net, int batch,
                       const cgm_fgmres_params* params, cgm_fgmres_result* results, double* dev_hist);

/* ---- tile preconditioner ------------------------------------------------------ */
/* No counterpart in the reference, which has no preconditioner
 * (B/CUDA_GMRES/GMRES.cuh:22-446).  The grid is cut into T x T tiles
 * (T = 8, 16, 32, 64 or 128, N % T == 0) and P applies the exact inverse of
 * the operator restricted to each tile,
 *     (P v)_t = (I - G_t diag(chi_t))^{-1} v_t,  G_t[p, q] = g[|p_r-q_r|, |p_c-q_c|],
 * with g the N x N generator block of the operator.
 * T <= 32: the inverses are formed once per map in FP32 complex (batched LU
 * with partial pivoting) and take T^4 * 8 bytes per tile: (N/T)^2 T^4 * 8 =
 * 512 MiB per map at N = 512, T = 16.
 * T = 64, 128 (reduced form): with S the cells of the tile where chi != 0,
 *     (P v)_t = v_t + G_t[:, S] W_t v_t[S],  W_t = (I - D_S G_SS)^{-1} D_S,
 * D_S = diag(chi_S).  W_t (|S| x |S|, |S| padded to 16) is formed once per map
 * by a symmetric block sweep of D_S^{-1} - G_SS in FP32 complex; it is complex
 * symmetric,
 * and the lower triangle of its 128 x 128 blocks is kept in FP16 scaled per
 * tile; G_t acting on W_t v_t[S] is a convolution done with FFTs of size 2T.
 * On the benchmark maps (3/8 of the cells with chi != 0) W takes 0.41 GB at
 * N = 512, T = 64 and 1.33 GB at T = 128; the factorisation holds an FP32
 * workspace of the full matrices while it runs.  The FP16
 * storage makes P about 1e-2 less exact on the near-resonant tiles of the
 * benchmark maps (1e-5 on low-contrast ones). */
typedef struct cgm_tiles* cgm_tiles_t;

/* The generator block g[|di|, |dj|] (N*N complex128, device memory) of the
 * benchmark operator (as cgm_spectrum_helmholtz) or of the reference's
 * operator (as cgm_spectrum_reference). */
int cgm_block_helmholtz(cgm_ctx_t ctx, double kh, cgm_complex128* dev_block);
int cgm_block_reference(cgm_ctx_t ctx, double wave_number, double row_shift, double col_shift,
                        cgm_complex128* dev_block);
/* Workspace for up to `maps` contrast maps at the context's N. */
int cgm_tiles_create(cgm_ctx_t ctx, int T, int maps, cgm_tiles_t* tiles);
int cgm_tiles_destroy(cgm_tiles_t tiles);
/* Device bytes the workspace holds (grows with the largest factorisation). */
int cgm_tiles_memory(cgm_tiles_t tiles, size_t* bytes);
/* Forms and inverts every tile's operator for `maps` maps (chi is
 * [maps][N*N] with chi_stride >= N*N, or one map with maps = 1).  One map
 * serves any number of right-hand sides.  Synchronous.  Returns
 * CGM_ERROR_NONFINITE if a tile's operator is singular (T <= 32) or a pivot
 * block of its reduced matrix is (T >= 64). */
int cgm_tiles_factor(cgm_tiles_t tiles, const cgm_complex128* dev_block, const float* dev_chi, long chi_stride,
                     int maps);
/* y[b] = P x[b] for b < batch ([batch][N*N]; map b, or the one map).
 * Asynchronous on the context stream.  With one map and several systems the
 * reduced form (T >= 64) runs the systems of each W block back to back on one
 * XCD, whose L2 serves the repeated reads (16 systems fetch W from HBM about
 * 1.5 times per call); the dense form (T <= 32) applies each inverse to four
 * systems per pass. */
int cgm_tiles_apply(cgm_tiles_t tiles, const cgm_complex64* dev_x, cgm_complex64* dev_y, int batch);
/* Algorithmic bytes of one cgm_tiles_apply over `batch` systems: the inverse
 * entries it streams (every entry of every tile in FP32 for T <= 32, once per
 * group of four right-hand sides of one map; the stored FP16 blocks of W for
 * T >= 64, once per system) plus 16 * N*N per system. */
int cgm_tiles_apply_bytes(cgm_tiles_t tiles, int batch, double* bytes);
/* Diagnostics: fills the inverses of one map with synthetic values (no
 * factorisation) and returns in *ms the average device time of `reps`
 * applications to one vector (HIP events on the context stream).  The
 * workspace must be factored again before it is used. */
int cgm_profile_tiles(cgm_tiles_t tiles, int reps, float* ms);

/* The right preconditioner of cgm_fgmres_pc / cgm_fgmres_pc_batched: the tile
 * inverses (tiles), the network (net), both, or the identity (both NULL).
 * Both set is the two-level preconditioner: z1 = T v, r = v - A z1,
 * P(v) = z1 + ||r|| F(r / ||r||), the network F a global correction on what
 * the tile inverses leave (trained on those residuals). */
typedef struct {
  cgm_tiles_t tiles;
  cgm_fno_t net;
} cgm_precond;

/* Precision of the long-recurrence flexible steps (restart > 64): with an FP64
 * spectrum set here (from cgm_spectrum_*(..., dev_gamma64)), w = A z_j and
 * each cycle's residual are formed by the complex128 FFT passes from the
 * complex64 vectors and rounded once.  A strong preconditioner returns z_j
 * of large norm, and the complex64 product's error, about 1e-7 ||A|| ||z_j||,
 * then stalls the true residual near 1e-5 while the estimates keep falling;
 * the FP64 product removes that floor.  NULL restores the complex64 passes.
 * The array must outlive the solves that use it.  No counterpart in the
 * reference. */
int cgm_set_flexible_fp64(cgm_ctx_t ctx, const cgm_complex128* dev_gamma64);

/* cgm_fgmres / cgm_fgmres_batched with the preconditioner given as a
 * cgm_precond.  cgm_fgmres(..., net, ...) is cgm_fgmres_pc with {NULL, net}.
 * The batched form needs the factored maps to match chi: one map (chi_stride
 * = 0) or one per system. */
int cgm_fgmres_pc(cgm_ctx_t ctx, const cgm_complex64* dev_gamma, const float* dev_chi,
                  const cgm_complex64* dev_b, cgm_complex64* dev_x, const cgm_precond* pc,
                  const cgm_fgmres_params* params, cgm_fgmres_result* result, double* dev_hist);
int cgm_fgmres_pc_batched(cgm_ctx_t ctx, const cgm_complex64* dev_gamma, const float* dev_chi, long chi_stride,
                          const cgm_complex64* dev_b, cgm_complex64* dev_x, const cgm_precond* pc, int batch,
                          const cgm_fgmres_params* params, cgm_fgmres_result* results, double* dev_hist);

/* cgm_gmres_refine with the correction solved by flexible GMRES with the
 * preconditioner pc (inner_restart up to 16384; above 64 the whole basis stays
 * in device memory).  The tiles of pc must have been factored for the map
 * chi rounded to float.  No counterpart in the reference. */
int cgm_gmres_refine_pc(cgm_ctx_t ctx, const cgm_complex64* dev_gamma,
                        const cgm_complex128* dev_gamma64, const double* dev_chi,
                        const cgm_complex128* dev_b, cgm_complex128* dev_x, const cgm_precond* pc,
                        const cgm_refine_params* params, cgm_refine_result* result);

/* ---- reference-compatible entry points ------------------------------------ */
/* The reference's plane wave exp(-i k i) with k = 2*3.14f/(N/6.f)
 * (B/CUDA_GMRES/GMRES_kernels.cuh:12,557-572), on the device. */
int cgm_reference_plane_wave(cgm_ctx_t ctx, cgm_complex64* dev_x);
/* CHI = (2.25 - 1) k^2 of B/CUDA_GMRES/GMRES_kernels.cuh:18-19, float. */
float cgm_reference_chi(int N);
/* Contrast used by the two entry points below for mask cells (default
 * cgm_reference_chi(N)). */
int cgm_set_mask_contrast(cgm_ctx_t ctx, float chi);
/* Number of timestamps Fast_GMRES_with_CUDA writes for maxiter
 * (B/CUDA_GMRES/launch_GMRES.cuh:245-307). */
unsigned cgm_n_timestamps(unsigned maxiter);

/* Drop-in for Fast_GMRES_with_CUDA (B/CUDA_GMRES/GMRES.cuh:22-35), same
 * parameter order and meaning:
 *   dev_gamma_array   spectrum from cgm_spectrum_* (caller owned)
 *   dev_mask          bool[N*N]; the contrast is cgm_set_mask_contrast * mask
 *   dev_solution      in: x0, out: x
 *   dev_actual_residual  callee-allocated device float[maxiter+1], caller
 *                     frees with cgm_free: [0] = ||r0||, [i] = residual
 *                     estimate after i steps
 *   GMRES_n           Krylov steps done
 *   plan              the context (in place of cufftHandle + cuBLAS handle)
 *   tolerance         absolute; no iteration when ||r0|| <= tolerance
 *   for_gradient      1: solve A^T x = e_{h_index_of_max}, else A x = b with
 *                     b the reference plane wave
 *   h_res_vs_tol_p    final estimate > tolerance
 *   maxiter           Krylov dimension (fixed, no restart)
 *   h_computation_times  caller array of cgm_n_timestamps(maxiter) entries,
 *                     stamped per phase (may be NULL)
 * The Givens rotations are unitary, so the estimates are the true GMRES
 * residual norms in exact arithmetic.  Returns a cgm_status. */
struct timespec;
int cgm_Fast_GMRES_with_CUDA(const cgm_complex64* dev_gamma_array, const bool* dev_mask,
                             cgm_complex64* dev_solution, float** dev_actual_residual,
                             unsigned int* GMRES_n, cgm_ctx_t plan, float tolerance,
                             bool for_gradient, unsigned int h_index_of_max,
                             bool* h_res_vs_tol_p, unsigned int maxiter,
                             struct timespec* h_computation_times, unsigned int N);

/* Drop-in for G_x_fft_matvec followed by _2D_to_1D_kernel
 * (B/CUDA_GMRES/Fast_matvec.cuh:1-28, GMRES_kernels.cuh:286-326): writes
 * dev_out = x - G(chi mask x) on the N x N grid (the (2N-1)^2 intermediate of
 * the reference stays internal). */
int cgm_G_x_fft_matvec(const cgm_complex64* dev_gamma_array, const bool* dev_mask,
                       const cgm_complex64* dev_solution, cgm_complex64* dev_out,
                       cgm_ctx_t plan, unsigned int N);
/* The gradient pairing: the unmasked G_x_fft_matvec overload
 * (B/CUDA_GMRES/Fast_matvec.cuh:31-56) followed by the masked crop
 * _2D_to_1D_kernel(dev_mask, ...) (B/CUDA_GMRES/GMRES_kernels.cuh:305-326),
 * i.e. dev_out = x - chi mask G(x) = A^T x. */
int cgm_G_x_fft_matvec_T(const cgm_complex64* dev_gamma_array, const bool* dev_mask,
                         const cgm_complex64* dev_solution, cgm_complex64* dev_out,
                         cgm_ctx_t plan, unsigned int N);

/* Drop-in for cuSolve_LES (B/CUDA_GMRES/GMRES.cuh:450-511): solves the m x m
 * column-major system A y = B in place (B <- y) by LU with partial pivoting,
 * on the device in FP64.  dev_work, dev_Ipiv and dev_info as in the reference
 * (dev_work may be NULL); *dev_info = 0 on success or k if U(k,k) = 0. */
int cgm_cuSolve_LES(cgm_complex64* dev_A, cgm_complex64* dev_B, int m, cgm_ctx_t ctx,
                    cgm_complex64* dev_work, int* dev_Ipiv, int* dev_info);

/* ---- diagnostics ------------------------------------------------------------ */
/* Average device time (ms, HIP events on the context stream, reps launches
 * each) of the kernels of one Arnoldi step at basis size restart/2 + 1:
 * ms[0..2] the three FFT passes of the matvec, ms[3] first CGS projection,
 * ms[4] update + second projection, ms[5] update + norm + Givens.  Uses
 * dev_x as every basis vector; overwrites the context's solver workspace. */
int cgm_profile_step(cgm_ctx_t ctx, const cgm_complex64* dev_gamma, const float* dev_chi,
                     const cgm_complex64* dev_x, int restart, int reps, float* ms);

/* Average device time (ms, HIP events on the context stream, reps launches
 * each) of the preconditioner's kernels on `batch` vectors ([batch][N*N] chi,
 * v, z): ms[0] lift with the first row projection, [1] stand-alone row
 * projection (backward pass), [2] x-DFT, [3] mode mixing, [4] inverse x-DFT,
 * [5] cell GEMM with GELU and the next row projection (layers 1-2), [6] last
 * cell GEMM with the projection, [7] a whole forward pass, [8] the cell GEMM
 * of layer 0, which lifts its row from chi and v itself (ms holds 9
 * entries).  Overwrites the network's activations. */
int cgm_fno_profile(cgm_fno_t net, const float* dev_chi, const cgm_complex64* dev_v,
                    cgm_complex64* dev_z, int batch, int reps, float* ms);

/* Average device time (ms, HIP events on the context stream, reps launches
 * each) of the long-recurrence step kernels (restart > 64) at basis size j:
 * ms[0] projection of u and w on the j basis vectors, [1] reduction of its
 * partial sums, [2] Givens update of the completed column (includes a 4-byte
 * memset per launch), [3] basis update forming q_j and u_{j+1} (ms holds 4
 * entries).  Overwrites the context's long-recurrence workspace. */
int cgm_profile_longrec(cgm_ctx_t ctx, int j, int reps, float* ms);
/* Live timing of the long-recurrence projection and update kernels inside
 * the solves that follow (HIP events around each launch on the context's
 * stream, read back every 32 steps; about one host synchronisation per 32
 * steps while on).  enable = 0 turns it off; either call zeroes the sums.
 * out[0..2] = projection: summed milliseconds, algorithmic bytes (8 B per
 * cell for each basis vector read and each vector read or written),
 * launches; out[3..5] = the same for the update; out[6..8] for the tile
 * inverses' streaming kernel (k_tile_apply for T <= 32, k_tile_ws for
 * T >= 64) with the bytes of cgm_tiles_apply_bytes. */
int cgm_longrec_timing(cgm_ctx_t ctx, int enable);
int cgm_longrec_timing_read(cgm_ctx_t ctx, double* out);

/* ---- memory and host helpers ---------------------------------------------- */
int cgm_malloc(void** ptr, size_t bytes);
int cgm_free(void* ptr);
int cgm_memcpy_h2d(void* dst, const void* src, size_t bytes);
int cgm_memcpy_d2h(void* dst, const void* src, size_t bytes);
int cgm_synchronize(cgm_ctx_t ctx);
/* FP64 J0, J1, Y0, Y1 (host); H_n^(1) = J_n + i Y_n. */
void cgm_bessel01(double x, double* j0, double* j1, double* y0, double* y1);
/* out[q] = (i/4) H0^(1)(k r[q]) on the host in FP64 (get_greenfun,
 * B/PyGMRES/GMRES.py:20-21); NaN where k r[q] <= 0. */
int cgm_green_values(const double* r, size_t n, double k, cgm_complex128* out);

#ifdef __cplusplus
}
#endif
#endif /* CUGMRES_H */
