// A reference-style caller of the solver, compiled with g++ against
// include/cugmres_compat.hpp and linked to cugmres_amd/libcugmres.so.
//
// It runs the sequence of launch_GMRES (B/CUDA_GMRES/launch_GMRES.cuh:93-233,
// B/ = the reference's benchmark folder) for one grid size with the
// reference's own names: plan and kernel spectrum from the host G_prep array,
// mask upload, x0 = the reference plane wave (init_x0_kernel), one
// Fast_GMRES_with_CUDA call, the residual read at [GMRES_n], the solution
// copied back, then the caller frees dev_actual_residual.  It then repeats
// the solve with for_gradient = true (A^T x = e_idx, GMRES.cuh:97-111) and
// exercises both G_x_fft_matvec overloads and cuSolve_LES.
//
// usage: caller N maxiter gprep.bin mask.bin out_prefix
//   gprep.bin  (2N-1)^2 complex64 (the reference's G_prep array)
//   mask.bin   N*N bytes (0/1)
// writes out_prefix + {solution,gradient,matvec,matvec_green,les}.bin and
// prints one line "GMRES_n=<n> residual=<r> res_vs_tol=<0|1> ...".
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "cugmres_compat.hpp"

static std::vector<char> read_file(const char* path, size_t bytes) {
  std::vector<char> buf(bytes);
  FILE* f = fopen(path, "rb");
  if (!f || fread(buf.data(), 1, bytes, f) != bytes) {
    fprintf(stderr, "cannot read %zu bytes from %s\n", bytes, path);
    exit(2);
  }
  fclose(f);
  return buf;
}

static void write_file(const std::string& path, const void* data, size_t bytes) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f || fwrite(data, 1, bytes, f) != bytes) {
    fprintf(stderr, "cannot write %s\n", path.c_str());
    exit(2);
  }
  fclose(f);
}

static void download(std::vector<cuComplex>& h, const cuComplex* d) {
  CGM_COMPAT_CALL(cgm_memcpy_d2h(h.data(), d, h.size() * sizeof(cuComplex)));
}

int main(int argc, char** argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: %s N maxiter gprep.bin mask.bin out_prefix\n", argv[0]);
    return 2;
  }
  const unsigned int N = (unsigned)atoi(argv[1]);
  const unsigned int maxiter = (unsigned)atoi(argv[2]);
  const std::string out = argv[5];
  const size_t n = (size_t)N * N;
  const size_t ng = (size_t)(2 * N - 1) * (2 * N - 1);
  std::vector<char> gprep = read_file(argv[3], ng * sizeof(cuComplex));
  std::vector<char> mask_bytes = read_file(argv[4], n);
  bool* h_mask = (bool*)malloc(n * sizeof(bool));
  for (size_t e = 0; e < n; ++e) h_mask[e] = mask_bytes[e] != 0;

  // launch_GMRES.cuh:95-101: plan and handles -> one context
  cufftHandle plan = create_plan(N);
  cublasHandle_t handle = plan;
  cusolverDnHandle_t cusolverH = plan;
  // launch_GMRES.cuh:115: timestamps sized for maxiter
  const unsigned nts = cgm_n_timestamps(maxiter);
  timespec* h_computation_times = (timespec*)malloc(nts * sizeof(timespec));

  bool* dev_mask = nullptr;
  cuComplex* dev_solution = nullptr;
  CGM_COMPAT_CALL(cgm_malloc((void**)&dev_mask, n * sizeof(bool)));
  CGM_COMPAT_CALL(cgm_malloc((void**)&dev_solution, n * sizeof(cuComplex)));
  CGM_COMPAT_CALL(cgm_memcpy_h2d(dev_mask, h_mask, n * sizeof(bool)));
  // launch_GMRES.cuh:133-142: G_prep upload + forward FFT -> kernel spectrum
  cuComplex* dev_gamma_array = nullptr;
  get_gamma_array(&dev_gamma_array, (const cuComplex*)gprep.data(), plan, N);

  // launch_GMRES.cuh:158: x0 = b (init_x0_kernel)
  CGM_COMPAT_CALL(cgm_reference_plane_wave(plan, dev_solution));
  CGM_COMPAT_CALL(cgm_synchronize(plan));
  memset(h_computation_times, 0, nts * sizeof(timespec));

  float* dev_actual_residual = nullptr;
  unsigned int GMRES_n = 0;
  bool h_res_vs_tol = true;
  const float tolerance = 0.001f;
  Fast_GMRES_with_CUDA((const cuComplex*)dev_gamma_array, (const bool*)dev_mask, (cuComplex*)dev_solution,
                       (float**)&dev_actual_residual, (unsigned int*)&GMRES_n, (cufftHandle)plan,
                       (cublasHandle_t*)&handle, tolerance, false, 0, (bool*)&h_res_vs_tol, maxiter,
                       (cusolverDnHandle_t)cusolverH, (timespec*)h_computation_times, N);
  // launch_GMRES.cuh:191: the caller reads the estimate after GMRES_n steps
  float residual = 0.f;
  CGM_COMPAT_CALL(cgm_memcpy_d2h(&residual, dev_actual_residual + GMRES_n, sizeof(float)));
  float residual0 = 0.f;
  CGM_COMPAT_CALL(cgm_memcpy_d2h(&residual0, dev_actual_residual, sizeof(float)));
  std::vector<cuComplex> h_sol(n);
  download(h_sol, dev_solution);
  write_file(out + "solution.bin", h_sol.data(), n * sizeof(cuComplex));
  long stamped = 0;
  for (unsigned t = 0; t < nts; ++t) stamped += h_computation_times[t].tv_sec != 0;
  // launch_GMRES.cuh:233: the caller frees the callee-allocated residual array
  CGM_COMPAT_CALL(cgm_free(dev_actual_residual));

  // the gradient solve: A^T x = e_{h_index_of_max}, x0 = the plane wave again
  const unsigned int h_index_of_max = (unsigned)(n / 2 + N / 3);
  CGM_COMPAT_CALL(cgm_reference_plane_wave(plan, dev_solution));
  float* dev_grad_residual = nullptr;
  unsigned int grad_n = 0;
  bool grad_vs_tol = true;
  Fast_GMRES_with_CUDA(dev_gamma_array, dev_mask, dev_solution, &dev_grad_residual, &grad_n, plan, &handle, 0.f,
                       true, h_index_of_max, &grad_vs_tol, maxiter, cusolverH, h_computation_times, N);
  download(h_sol, dev_solution);
  write_file(out + "gradient.bin", h_sol.data(), n * sizeof(cuComplex));
  float grad_res = 0.f;
  CGM_COMPAT_CALL(cgm_memcpy_d2h(&grad_res, dev_grad_residual + grad_n, sizeof(float)));
  CGM_COMPAT_CALL(cgm_free(dev_grad_residual));

  // both G_x_fft_matvec overloads on x = the plane wave
  cuComplex* dev_x = nullptr;
  cuComplex* dev_y = nullptr;
  CGM_COMPAT_CALL(cgm_malloc((void**)&dev_x, n * sizeof(cuComplex)));
  CGM_COMPAT_CALL(cgm_malloc((void**)&dev_y, n * sizeof(cuComplex)));
  CGM_COMPAT_CALL(cgm_reference_plane_wave(plan, dev_x));
  G_x_fft_matvec(dev_gamma_array, dev_mask, dev_x, dev_y, plan, N);
  download(h_sol, dev_y);
  write_file(out + "matvec.bin", h_sol.data(), n * sizeof(cuComplex));
  G_x_fft_matvec(dev_gamma_array, dev_x, dev_y, plan, N);
  download(h_sol, dev_y);
  write_file(out + "matvec_green.bin", h_sol.data(), n * sizeof(cuComplex));

  // cuSolve_LES on a fixed 4 x 4 column-major system
  const int m = 4;
  cuComplex hA[m * m], hB[m];
  for (int r = 0; r < m; ++r) {
    hB[r].re = 1.0f + r;
    hB[r].im = -0.5f * r;
    for (int c = 0; c < m; ++c) {
      hA[r + c * m].re = (r == c ? 4.0f : 0.0f) + 0.25f * (float)((r + 2 * c) % 3);
      hA[r + c * m].im = 0.125f * (float)(r - c);
    }
  }
  cuComplex *dA = nullptr, *dB = nullptr;
  int *dIpiv = nullptr, *dInfo = nullptr;
  CGM_COMPAT_CALL(cgm_malloc((void**)&dA, sizeof(hA)));
  CGM_COMPAT_CALL(cgm_malloc((void**)&dB, sizeof(hB)));
  CGM_COMPAT_CALL(cgm_malloc((void**)&dIpiv, m * sizeof(int)));
  CGM_COMPAT_CALL(cgm_malloc((void**)&dInfo, sizeof(int)));
  CGM_COMPAT_CALL(cgm_memcpy_h2d(dA, hA, sizeof(hA)));
  CGM_COMPAT_CALL(cgm_memcpy_h2d(dB, hB, sizeof(hB)));
  unsigned int clock_i = 0;
  timespec les_times[2];
  cuSolve_LES(dA, dB, m, cusolverH, nullptr, dIpiv, dInfo, les_times, &clock_i);
  CGM_COMPAT_CALL(cgm_memcpy_d2h(hB, dB, sizeof(hB)));
  write_file(out + "les.bin", hB, sizeof(hB));

  printf("GMRES_n=%u residual0=%.9g residual=%.9g res_vs_tol=%d stamped=%ld nts=%u grad_n=%u grad_residual=%.9g "
         "clock_i=%u\n",
         GMRES_n, residual0, residual, (int)h_res_vs_tol, stamped, nts, grad_n, grad_res, clock_i);

  CGM_COMPAT_CALL(cgm_free(dA));
  CGM_COMPAT_CALL(cgm_free(dB));
  CGM_COMPAT_CALL(cgm_free(dIpiv));
  CGM_COMPAT_CALL(cgm_free(dInfo));
  CGM_COMPAT_CALL(cgm_free(dev_x));
  CGM_COMPAT_CALL(cgm_free(dev_y));
  CGM_COMPAT_CALL(cgm_free(dev_mask));
  CGM_COMPAT_CALL(cgm_free(dev_solution));
  CGM_COMPAT_CALL(cgm_free(dev_gamma_array));
  CGM_COMPAT_CALL(cgm_destroy(plan));
  free(h_computation_times);
  free(h_mask);
  return 0;
}
