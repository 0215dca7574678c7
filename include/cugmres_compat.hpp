/*
 * cugmres_compat.hpp -- the reference's C++ entry points under their own
 * names, forwarding to the C ABI of cugmres.h.
 *
 * A caller written against the reference's headers (B/CUDA_GMRES/GMRES.cuh,
 * B/CUDA_GMRES/Fast_matvec.cuh; B/ = "Benchmarks of CUDA-based GMRES Solver
 * for Toeplitz and Hankel Matrices and Applications to Topology Optimization
 * of Photonic Devices/") includes this header instead, and its calls of
 *
 *   Fast_GMRES_with_CUDA   (B/CUDA_GMRES/GMRES.cuh:22-35)
 *   G_x_fft_matvec x 2     (B/CUDA_GMRES/Fast_matvec.cuh:1-6, 31-35)
 *   cuSolve_LES            (B/CUDA_GMRES/GMRES.cuh:450-458)
 *
 * compile unchanged: same names, parameter order, parameter meaning and
 * void return.  Error behaviour is the reference's as well: a failing call
 * prints file, line and reason to stderr and exits with EXIT_FAILURE (the
 * cudacall / cublascall macros of B/CUDA_GMRES/GMRES_kernels.cuh:21-94).
 *
 * The vendor handle types become the library context.  One cgm_ctx_t holds
 * the FFT plan, the BLAS and solver workspaces and the stream, so the caller
 * passes the same context wherever the reference passes a cufftHandle, a
 * cublasHandle_t or a cusolverDnHandle_t.  What the caller replaces around
 * the calls (handle creation, the forward FFT of G_prep, cudaMalloc) is
 * listed with the deviations in cugmres.h; the helpers at the end of this
 * file cover the first two.
 *
 * Deviations visible through these signatures:
 *   - dev_gamma_array is the library's kernel spectrum (cgm_spectrum_*),
 *     not the (2N-1)^2 cuFFT array;
 *   - G_x_fft_matvec writes the N x N result with crop, 1/(2N-1)^2 scale and
 *     mask applied (x - G(chi mask x), or G x for the unmasked overload);
 *     the reference writes the (2N-1)^2 intermediate and crops it with
 *     _2D_to_1D_kernel (B/CUDA_GMRES/GMRES_kernels.cuh:286-326);
 *   - *dev_actual_residual is released with cgm_free (or hipFree).
 */
#ifndef CUGMRES_COMPAT_HPP
#define CUGMRES_COMPAT_HPP

#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "cugmres.h"

#ifndef CGM_COMPAT_NO_VENDOR_NAMES
/* The reference's vendor types, as the library's own. */
typedef cgm_complex64 cuComplex;
typedef cgm_ctx_t cufftHandle;
typedef cgm_ctx_t cublasHandle_t;
typedef cgm_ctx_t cusolverDnHandle_t;
#endif

namespace cgm_compat {

inline void fail(int status, const char* call, const char* file, int line) {
  fprintf(stderr, "cuGMRES error:\nFile = %s\nLine = %d\nCall = %s\nReason = %s: %s\n", file, line, call,
          cgm_status_string(status), cgm_last_error());
  exit(EXIT_FAILURE);
}

}  // namespace cgm_compat

#define CGM_COMPAT_CALL(call)                                                  \
  do {                                                                         \
    const int cgm_st_ = (call);                                                \
    if (cgm_st_ != CGM_SUCCESS) cgm_compat::fail(cgm_st_, #call, __FILE__, __LINE__); \
  } while (0)

/* B/CUDA_GMRES/GMRES.cuh:22-35.  handle_p and cusolverH are accepted for the
 * signature; the work runs on `plan`, the context. */
inline void Fast_GMRES_with_CUDA(const cgm_complex64* dev_gamma_array, const bool* dev_mask,
                                 cgm_complex64* dev_solution, float** dev_actual_residual,
                                 unsigned int* GMRES_n, cgm_ctx_t plan, cgm_ctx_t* handle_p,
                                 const float tolerance, const bool for_gradient,
                                 const unsigned int h_index_of_max, bool* h_res_vs_tol_p,
                                 unsigned int maxiter, cgm_ctx_t cusolverH,
                                 struct timespec* h_computation_times, const unsigned int N) {
  (void)handle_p;
  (void)cusolverH;
  CGM_COMPAT_CALL(cgm_Fast_GMRES_with_CUDA(dev_gamma_array, dev_mask, dev_solution, dev_actual_residual, GMRES_n,
                                           plan, tolerance, for_gradient, h_index_of_max, h_res_vs_tol_p, maxiter,
                                           h_computation_times, N));
}

/* B/CUDA_GMRES/Fast_matvec.cuh:1-6: the masked product, here written as
 * dev_matmul_out = x - G(chi mask x) on the N x N grid. */
inline void G_x_fft_matvec(cgm_complex64* dev_gamma_array, bool* dev_mask, cgm_complex64* dev_solution,
                           cgm_complex64* dev_matmul_out, cgm_ctx_t plan, const unsigned int N) {
  CGM_COMPAT_CALL(cgm_G_x_fft_matvec(dev_gamma_array, dev_mask, dev_solution, dev_matmul_out, plan, N));
}

/* B/CUDA_GMRES/Fast_matvec.cuh:31-35: the unmasked product of the gradient
 * path, here dev_matmul_out = G x on the N x N grid (the reference crops it
 * with the mask afterwards; cgm_G_x_fft_matvec_T does both). */
inline void G_x_fft_matvec(cgm_complex64* dev_gamma_array, cgm_complex64* dev_solution,
                           cgm_complex64* dev_matmul_out, cgm_ctx_t plan, const unsigned int N) {
  (void)N;
  CGM_COMPAT_CALL(cgm_matvec(plan, dev_gamma_array, nullptr, dev_solution, dev_matmul_out, CGM_OP_GREEN));
  CGM_COMPAT_CALL(cgm_synchronize(plan));
}

/* B/CUDA_GMRES/GMRES.cuh:450-458: LU solve of the column-major m x m system
 * (dev_B <- solution), stamping two entries of h_computation_times at
 * *clock_i_p as the reference does (GMRES.cuh:469,511).  A nonzero *dev_info
 * exits, as the reference's check after getrf does (GMRES.cuh:484-488). */
inline void cuSolve_LES(cgm_complex64* dev_A, cgm_complex64* dev_B, const int m, cgm_ctx_t cusolverH,
                        cgm_complex64* dev_work, int* dev_Ipiv, int* dev_info,
                        struct timespec* h_computation_times, unsigned int* clock_i_p) {
  if (h_computation_times && clock_i_p) clock_gettime(CLOCK_REALTIME, h_computation_times + (*clock_i_p)++);
  CGM_COMPAT_CALL(cgm_cuSolve_LES(dev_A, dev_B, m, cusolverH, dev_work, dev_Ipiv, dev_info));
  int h_info = 0;
  CGM_COMPAT_CALL(cgm_memcpy_d2h(&h_info, dev_info, sizeof(int)));
  if (h_info != 0) {
    printf("cuSolve_LES: LU factorisation failed!\tinfo = %i\n", h_info);
    exit(1);
  }
  if (h_computation_times && clock_i_p) clock_gettime(CLOCK_REALTIME, h_computation_times + (*clock_i_p)++);
}

/* What launch_GMRES does around the calls (B/CUDA_GMRES/launch_GMRES.cuh:95-142),
 * in the library's terms: the context in place of cufftPlan2d + cublasCreate
 * + cusolverDnCreate, and the kernel spectrum from the host G_prep array in
 * place of cudaMemcpy + cufftExecC2C(FORWARD).  *dev_gamma_array is
 * allocated here and released with cgm_free. */
inline cgm_ctx_t create_plan(unsigned int N, int device = 0) {
  cgm_ctx_t c = nullptr;
  CGM_COMPAT_CALL(cgm_create(&c, (int)N, device));
  return c;
}

inline void get_gamma_array(cgm_complex64** dev_gamma_array, const cgm_complex64* h_gprep, cgm_ctx_t plan,
                            unsigned int N) {
  size_t len = 0;
  CGM_COMPAT_CALL(cgm_spectrum_size(plan, &len));
  const size_t ng = (size_t)(2 * N - 1) * (2 * N - 1);
  void* gp = nullptr;
  CGM_COMPAT_CALL(cgm_malloc(&gp, ng * sizeof(cgm_complex64)));
  CGM_COMPAT_CALL(cgm_memcpy_h2d(gp, h_gprep, ng * sizeof(cgm_complex64)));
  void* g = nullptr;
  CGM_COMPAT_CALL(cgm_malloc(&g, len * sizeof(cgm_complex64)));
  CGM_COMPAT_CALL(cgm_spectrum_from_gprep(plan, (const cgm_complex64*)gp, (cgm_complex64*)g));
  CGM_COMPAT_CALL(cgm_synchronize(plan));
  CGM_COMPAT_CALL(cgm_free(gp));
  *dev_gamma_array = (cgm_complex64*)g;
}

#endif /* CUGMRES_COMPAT_HPP */
