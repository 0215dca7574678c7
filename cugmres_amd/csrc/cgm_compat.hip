// Entry points with the reference's names, parameter order and meaning
// (B/CUDA_GMRES/GMRES.cuh:22-35,450-458; B/CUDA_GMRES/Fast_matvec.cuh:1-6,31-35),
// built on the library's own operator and solver.  Deviations are listed in
// include/cugmres.h: status codes instead of exit(), a library context in
// place of cuFFT/cuBLAS/cuSOLVER handles, unitary Givens rotations.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <time.h>

#include "cgm_internal.h"

using namespace cgm;

// B/CUDA_GMRES/GMRES_kernels.cuh:12,18-19: (EPSILON-1)*WAVE_NUMBER*WAVE_NUMBER
// expands to a left-to-right float product with WAVE_NUMBER = 2*3.14f/(N/6.f).
extern "C" float cgm_reference_chi(int N) {
  volatile float lam = (float)N / 6.0f;
  volatile float v = (2.25f - 1.0f) * 2.0f;
  v = v * 3.14f;
  v = v / lam;
  v = v * 2.0f;
  v = v * 3.14f;
  v = v / lam;
  return v;
}

extern "C" int cgm_set_mask_contrast(cgm_ctx_t c, float chi) {
  CGM_CHECK(c, CGM_ERROR_INVALID_VALUE, "null context");
  c->mask_chi = chi;
  return CGM_SUCCESS;
}

// B/CUDA_GMRES/launch_GMRES.cuh:245-307
extern "C" unsigned cgm_n_timestamps(unsigned maxiter) {
  unsigned n = 1 + 6 + 11;
  unsigned gi;
  for (gi = 1; gi < maxiter; ++gi) n += 3 + 2 * (gi + 1) + 7;
  n += 4;
  n += gi;
  n += 2;
  return n;
}

__global__ void k_mask_chi(const bool* __restrict__ mask, float chi, float* __restrict__ out, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    out[e] = mask[e] ? chi : 0.0f;
}

// init_x0_kernel (B/CUDA_GMRES/GMRES_kernels.cuh:557-572), ALPHA = 0: the
// phase is the double product -WAVE_NUMBER * i rounded to float.
__global__ void k_ref_plane_wave(float2* __restrict__ x, int N, float k) {
  const long n = (long)N * N;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / N);
    const float ph = (float)(-(double)k * (double)i);
    float s, c;
    sincosf(ph, &s, &c);
    x[e] = make_float2(c, s);
  }
}

__global__ void k_unit(float2* __restrict__ x, long n, long idx) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    x[e] = make_float2(e == idx ? 1.0f : 0.0f, 0.0f);
}

__global__ void k_hist_to_float(const double* __restrict__ h, float* __restrict__ o, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) o[e] = (float)h[e];
}

static float ref_wave_number(int N) {
  volatile float lam = (float)N / 6.0f;
  volatile float v = 2.0f * 3.14f;
  return v / lam;
}

extern "C" int cgm_reference_plane_wave(cgm_ctx_t c, cgm_complex64* x) {
  CGM_CHECK(c && x, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_HIP(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_ref_plane_wave, dim3(1024), dim3(256), 0, c->stream, (float2*)x, c->N,
                     ref_wave_number(c->N));
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

static int ensure_tmp(cgm_context* c) {
  const long n = (long)c->N * c->N;
  if (!c->chi_tmp) CGM_HIP(hipMalloc((void**)&c->chi_tmp, sizeof(float) * n));
  if (!c->vec_tmp) CGM_HIP(hipMalloc((void**)&c->vec_tmp, sizeof(float2) * n));
  return CGM_SUCCESS;
}

static int mask_to_chi(cgm_context* c, const bool* mask) {
  const long n = (long)c->N * c->N;
  hipLaunchKernelGGL(k_mask_chi, dim3(1024), dim3(256), 0, c->stream, mask, c->mask_chi, c->chi_tmp, n);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

static void stamp(struct timespec* t, unsigned idx, unsigned n) {
  if (t && idx < n) clock_gettime(CLOCK_REALTIME, t + idx);
}

extern "C" int cgm_Fast_GMRES_with_CUDA(const cgm_complex64* gamma, const bool* mask,
                                        cgm_complex64* sol, float** dev_actual_residual,
                                        unsigned int* GMRES_n, cgm_ctx_t c, float tolerance,
                                        bool for_gradient, unsigned int h_index_of_max,
                                        bool* h_res_vs_tol_p, unsigned int maxiter,
                                        struct timespec* times, unsigned int N) {
  const unsigned nts = cgm_n_timestamps(maxiter);
  stamp(times, 0, nts);
  CGM_CHECK(c && gamma && mask && sol && dev_actual_residual && GMRES_n && h_res_vs_tol_p,
            CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK((int)N == c->N, CGM_ERROR_INVALID_VALUE, "N=%u does not match the context (%d)", N, c->N);
  CGM_CHECK(maxiter >= 1 && maxiter <= (unsigned)kMaxRestart, CGM_ERROR_INVALID_VALUE,
            "maxiter %u outside [1, %d]", maxiter, kMaxRestart);
  CGM_CHECK(!for_gradient || (long)h_index_of_max < (long)N * N, CGM_ERROR_INVALID_VALUE,
            "h_index_of_max out of range");
  CGM_HIP(hipSetDevice(c->device));
  const long n = (long)N * N;
  int rc = ensure_tmp(c);
  if (rc) return rc;
  // Every exit after the first allocation goes through `fail`, which frees
  // both device arrays; only a successful call hands `resid` to the caller.
  float* resid = nullptr;
  double* hist = nullptr;
  hipError_t he = hipSuccess;
  cgm_gmres_result r;
  unsigned steps = 0;
  double final_est = 0.0;
  if ((he = hipMalloc((void**)&resid, sizeof(float) * (maxiter + 1))) != hipSuccess) goto hip_fail;
  if ((he = hipMalloc((void**)&hist, sizeof(double) * (maxiter + 2))) != hipSuccess) goto hip_fail;
  if ((he = hipMemsetAsync(hist, 0, sizeof(double) * (maxiter + 2), c->stream)) != hipSuccess) goto hip_fail;
  if ((rc = mask_to_chi(c, mask))) goto fail;
  if (for_gradient)
    hipLaunchKernelGGL(k_unit, dim3(1024), dim3(256), 0, c->stream, (float2*)c->vec_tmp, n, (long)h_index_of_max);
  else
    hipLaunchKernelGGL(k_ref_plane_wave, dim3(1024), dim3(256), 0, c->stream, c->vec_tmp, (int)N,
                       ref_wave_number((int)N));
  if ((he = hipGetLastError()) != hipSuccess) goto hip_fail;
  stamp(times, 1, nts);
  // The reference iterates only when ||r0|| > tolerance (GMRES.cuh:153) and
  // then runs exactly maxiter Arnoldi steps (its early exit is commented out,
  // GMRES.cuh:253): one cycle with a zero tolerance inside it.
  rc = gmres_fixed_cycle(c, (const float2*)gamma, c->chi_tmp, c->vec_tmp, (float2*)sol, (int)maxiter,
                         for_gradient ? 1 : 0, (double)tolerance, hist, &r);
  if (rc) goto fail;
  steps = (unsigned)r.iterations;
  final_est = r.beta0;
  if (steps > 0 &&
      (he = hipMemcpy(&final_est, hist + steps, sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
    goto hip_fail;
  hipLaunchKernelGGL(k_hist_to_float, dim3((maxiter + 1 + 255) / 256), dim3(256), 0, c->stream,
                     hist, resid, (int)maxiter + 1);
  if ((he = hipGetLastError()) != hipSuccess) goto hip_fail;
  if ((he = hipStreamSynchronize(c->stream)) != hipSuccess) goto hip_fail;
  for (unsigned t = 2; t + 2 < nts; ++t) stamp(times, t, nts);
  *dev_actual_residual = resid;
  *GMRES_n = steps;
  *h_res_vs_tol_p = final_est > tolerance;
  stamp(times, nts - 2, nts);
  (void)hipFree(hist);
  stamp(times, nts - 1, nts);
  return CGM_SUCCESS;
hip_fail:
  cgm_set_last_error("HIP error %s in cgm_Fast_GMRES_with_CUDA: %s", hipGetErrorName(he), hipGetErrorString(he));
  rc = CGM_ERROR_DEVICE;
fail:
  (void)hipStreamSynchronize(c->stream);
  if (resid) (void)hipFree(resid);
  if (hist) (void)hipFree(hist);
  return rc;
}

static int gx_common(const cgm_complex64* gamma, const bool* mask, const cgm_complex64* x,
                     cgm_complex64* out, cgm_ctx_t c, unsigned N, int mode) {
  CGM_CHECK(c && gamma && mask && x && out, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK((int)N == c->N, CGM_ERROR_INVALID_VALUE, "N=%u does not match the context (%d)", N, c->N);
  CGM_HIP(hipSetDevice(c->device));
  int rc = ensure_tmp(c);
  if (!rc) rc = mask_to_chi(c, mask);
  if (!rc) rc = cgm_matvec(c, gamma, c->chi_tmp, x, out, mode);
  if (!rc) CGM_HIP(hipStreamSynchronize(c->stream));
  return rc;
}

extern "C" int cgm_G_x_fft_matvec(const cgm_complex64* gamma, const bool* mask,
                                  const cgm_complex64* x, cgm_complex64* out, cgm_ctx_t c,
                                  unsigned int N) {
  return gx_common(gamma, mask, x, out, c, N, CGM_OP_FORWARD);
}

extern "C" int cgm_G_x_fft_matvec_T(const cgm_complex64* gamma, const bool* mask,
                                    const cgm_complex64* x, cgm_complex64* out, cgm_ctx_t c,
                                    unsigned int N) {
  return gx_common(gamma, mask, x, out, c, N, CGM_OP_ADJOINT);
}

// ---------------------------------------------------------------------------
// cuSolve_LES: LU with partial pivoting (FP64) of a column-major m x m system.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_lu_solve(float2* __restrict__ A, float2* __restrict__ B, int m, double2* __restrict__ W,
           int* __restrict__ ipiv, int* __restrict__ info) {
  // W: m*m FP64 copy, column-major (W[r + c*m])
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) W[e] = make_double2(A[e].x, A[e].y);
  __shared__ double2 rhs[kMaxRestart];
  __shared__ double sval[256];
  __shared__ int sidx[256];
  __shared__ int s_info;
  for (int e = threadIdx.x; e < m; e += blockDim.x) rhs[e] = make_double2(B[e].x, B[e].y);
  if (threadIdx.x == 0) s_info = 0;
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    // pivot search |W[r, k]|, r >= k (ties to the smallest index)
    double best = -1.0;
    int bi = k;
    for (int r = k + threadIdx.x; r < m; r += blockDim.x) {
      const double2 v = W[r + (long)k * m];
      const double a = fabs(v.x) + fabs(v.y);
      if (a > best) { best = a; bi = r; }
    }
    sval[threadIdx.x] = best;
    sidx[threadIdx.x] = bi;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) {
        const double o = sval[threadIdx.x + s];
        const int oi = sidx[threadIdx.x + s];
        if (o > sval[threadIdx.x] || (o == sval[threadIdx.x] && oi < sidx[threadIdx.x])) {
          sval[threadIdx.x] = o;
          sidx[threadIdx.x] = oi;
        }
      }
      __syncthreads();
    }
    const int p = sidx[0];
    const double pv = sval[0];
    if (threadIdx.x == 0) {
      ipiv[k] = p + 1;
      if (pv == 0.0 && s_info == 0) s_info = k + 1;
    }
    if (p != k) {
      for (int c = threadIdx.x; c < m; c += blockDim.x) {
        const double2 t = W[k + (long)c * m];
        W[k + (long)c * m] = W[p + (long)c * m];
        W[p + (long)c * m] = t;
      }
      if (threadIdx.x == 0) { const double2 t = rhs[k]; rhs[k] = rhs[p]; rhs[p] = t; }
    }
    __syncthreads();
    const double2 d = W[k + (long)k * m];
    const double den = d.x * d.x + d.y * d.y;
    if (den > 0.0) {
      for (int r = k + 1 + threadIdx.x; r < m; r += blockDim.x) {
        const double2 a = W[r + (long)k * m];
        W[r + (long)k * m] = make_double2((a.x * d.x + a.y * d.y) / den, (a.y * d.x - a.x * d.y) / den);
      }
    }
    __syncthreads();
    const int rem = m - k - 1;
    for (int e = threadIdx.x; e < rem * rem; e += blockDim.x) {
      const int r = k + 1 + e % rem, c = k + 1 + e / rem;
      const double2 l = W[r + (long)k * m], u = W[k + (long)c * m];
      double2 w = W[r + (long)c * m];
      w.x -= l.x * u.x - l.y * u.y;
      w.y -= l.x * u.y + l.y * u.x;
      W[r + (long)c * m] = w;
    }
    if (threadIdx.x == 0) {
      for (int r = k + 1; r < m; ++r) {
        const double2 l = W[r + (long)k * m], b = rhs[k];
        rhs[r].x -= l.x * b.x - l.y * b.y;
        rhs[r].y -= l.x * b.y + l.y * b.x;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (s_info == 0) {
      for (int r = m - 1; r >= 0; --r) {
        double2 s = rhs[r];
        for (int c = r + 1; c < m; ++c) {
          const double2 u = W[r + (long)c * m], y = rhs[c];
          s.x -= u.x * y.x - u.y * y.y;
          s.y -= u.x * y.y + u.y * y.x;
        }
        const double2 d = W[r + (long)r * m];
        const double den = d.x * d.x + d.y * d.y;
        rhs[r] = make_double2((s.x * d.x + s.y * d.y) / den, (s.y * d.x - s.x * d.y) / den);
      }
    }
    *info = s_info;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < m * m; e += blockDim.x) A[e] = make_float2((float)W[e].x, (float)W[e].y);
  if (s_info == 0)
    for (int e = threadIdx.x; e < m; e += blockDim.x) B[e] = make_float2((float)rhs[e].x, (float)rhs[e].y);
}

extern "C" int cgm_cuSolve_LES(cgm_complex64* A, cgm_complex64* B, int m, cgm_ctx_t c,
                               cgm_complex64* work, int* ipiv, int* info) {
  (void)work;
  CGM_CHECK(c && A && B && ipiv && info, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(m >= 1 && m <= kMaxRestart, CGM_ERROR_INVALID_VALUE, "m=%d outside [1, %d]", m, kMaxRestart);
  CGM_HIP(hipSetDevice(c->device));
  // the FP64 copy of A lives in the context, allocated once at its largest
  // size (kMaxRestart^2), so repeated calls allocate nothing
  if (!c->les_ws) CGM_HIP(hipMalloc((void**)&c->les_ws, sizeof(double2) * (size_t)kMaxRestart * kMaxRestart));
  hipLaunchKernelGGL(k_lu_solve, dim3(1), dim3(256), 0, c->stream, (float2*)A, (float2*)B, m, c->les_ws, ipiv,
                     info);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  CGM_HIP(e);
  return CGM_SUCCESS;
}
