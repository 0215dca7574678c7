// C ABI of the solver: contexts, spectra, operator, restarted GMRES, FP64
// refinement and the reference-compatible entry points.  Host code; every
// device operation is enqueued on the context stream.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <vector>

#include "cgm_internal.h"
#include "cgm_special.h"

using namespace cgm;

static thread_local char g_last_error[512] = "";

void cgm_set_last_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

extern "C" const char* cgm_last_error(void) { return g_last_error; }

extern "C" const char* cgm_status_string(int s) {
  switch (s) {
    case CGM_SUCCESS: return "success";
    case CGM_ERROR_INVALID_VALUE: return "invalid value";
    case CGM_ERROR_DEVICE: return "device error";
    case CGM_ERROR_ALLOC: return "allocation failed";
    case CGM_ERROR_NOT_CONVERGED: return "not converged";
    case CGM_ERROR_NONFINITE: return "non-finite value";
    case CGM_ERROR_UNSUPPORTED: return "unsupported size";
    case CGM_ERROR_NOT_SYMMETRIC: return "kernel array not even-symmetric";
  }
  return "unknown status";
}

extern "C" void cgm_version(int* major, int* minor) {
  if (major) *major = CGM_VERSION_MAJOR;
  if (minor) *minor = CGM_VERSION_MINOR;
}

extern "C" void cgm_bessel01(double x, double* j0, double* j1, double* y0, double* y1) {
  double a, b, c, d;
  bessel01(x, a, b, c, d);
  if (j0) *j0 = a;
  if (j1) *j1 = b;
  if (y0) *y0 = c;
  if (y1) *y1 = d;
}

extern "C" int cgm_green_values(const double* r, size_t n, double k, cgm_complex128* out) {
  CGM_CHECK((r && out) || n == 0, CGM_ERROR_INVALID_VALUE, "null argument");
  for (size_t q = 0; q < n; ++q) {
    const double x = k * r[q];
    if (!(x > 0.0) || !isfinite(x)) {
      out[q].re = out[q].im = NAN;  // H0 is singular at 0
      continue;
    }
    double j0, j1, y0, y1;
    bessel01(x, j0, j1, y0, y1);
    out[q].re = -0.25 * y0;  // (i/4)(J0 + i Y0)
    out[q].im = 0.25 * j0;
  }
  return CGM_SUCCESS;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static int upload_twiddles(cgm_context* c) {
  const int M = c->M;
  std::vector<float2> t32(M);
  std::vector<double2> t64(M);
  for (int q = 0; q < M; ++q) {
    const double a = -2.0 * CGM_PI * (double)q / (double)M;
    t64[q] = make_double2(cos(a), sin(a));
    t32[q] = make_float2((float)t64[q].x, (float)t64[q].y);
  }
  CGM_HIP(hipMalloc((void**)&c->tw32, sizeof(float2) * M));
  CGM_HIP(hipMalloc((void**)&c->tw64, sizeof(double2) * M));
  CGM_HIP(hipMemcpy(c->tw32, t32.data(), sizeof(float2) * M, hipMemcpyHostToDevice));
  CGM_HIP(hipMemcpy(c->tw64, t64.data(), sizeof(double2) * M, hipMemcpyHostToDevice));
  return CGM_SUCCESS;
}

extern "C" int cgm_destroy(cgm_ctx_t c) {
  if (!c) return CGM_SUCCESS;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  krylov_free(c->ws);
  lr_free(c->lr);
  if (c->lrt.created)
    for (hipEvent_t e : c->lrt.ev) (void)hipEventDestroy(e);
  void* ptrs[] = {c->tw32, c->tw64, c->T32, c->T64, c->st, c->chi_tmp, c->vec_tmp, c->v64a, c->chi32, c->zb,
                  c->stB, c->T32B, c->les_ws, c->pc_v, c->pc_r, c->pc_z, c->chi64, c->z64, c->w64, c->T64B};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->hst) (void)hipHostFree(c->hst);
  if (c->hstB) (void)hipHostFree(c->hstB);
  if (c->gstream) (void)hipStreamDestroy(c->gstream);
  delete c;
  return CGM_SUCCESS;
}

extern "C" int cgm_create(cgm_ctx_t* out, int N, int device) {
  CGM_CHECK(out != nullptr, CGM_ERROR_INVALID_VALUE, "cgm_create: null output");
  *out = nullptr;
  CGM_CHECK(N >= 1 && N <= 2048, CGM_ERROR_UNSUPPORTED, "cgm_create: N=%d outside [1, 2048]", N);
  int ndev = 0;
  CGM_HIP(hipGetDeviceCount(&ndev));
  CGM_CHECK(device >= 0 && device < ndev, CGM_ERROR_INVALID_VALUE, "cgm_create: device %d of %d",
            device, ndev);
  CGM_HIP(hipSetDevice(device));
  {
    const int rs = noalloc_init(device);
    if (rs) return rs;
  }
  cgm_context* c = new cgm_context();
  c->N = N;
  c->device = device;
  int rc = plan_init(c->plan, N);
  if (rc) { delete c; cgm_set_last_error("cgm_create: no embedding for N=%d", N); return rc; }
  c->M = c->plan.M;
  c->plan64 = c->plan;
  c->mask_chi = cgm_reference_chi(N);
  if ((rc = upload_twiddles(c))) { cgm_destroy(c); return rc; }
  const size_t nm = (size_t)N * c->M;
  if (hipMalloc((void**)&c->T32, sizeof(float2) * nm) != hipSuccess ||
      hipMalloc((void**)&c->st, sizeof(DevState)) != hipSuccess ||
      hipHostMalloc((void**)&c->hst, sizeof(DevState), hipHostMallocDefault) != hipSuccess) {
    cgm_destroy(c);
    cgm_set_last_error("cgm_create: device allocation failed");
    return CGM_ERROR_ALLOC;
  }
  CGM_HIP(hipMemset(c->st, 0, sizeof(DevState)));
  if ((rc = cgm_autotune(c))) { cgm_destroy(c); return rc; }
  *out = c;
  return CGM_SUCCESS;
}

// ---------------------------------------------------------------------------
// autotuner: rows per work-group of passes 1/3 and columns per work-group of
// pass 2, timed with HIP events on zero data (the kernels' cost does not
// depend on the values)
// ---------------------------------------------------------------------------
extern "C" int cgm_autotune(cgm_ctx_t c) {
  CGM_CHECK(c, CGM_ERROR_INVALID_VALUE, "null context");
  CGM_HIP(hipSetDevice(c->device));
  const int N = c->N, M = c->M;
  const int nt = fft_threads(M);
  const long n = (long)N * N;
  const size_t H = (size_t)M / 2 + 1;
  float2 *x = nullptr, *gq = nullptr;
  float* chi = nullptr;
  CGM_HIP(hipMalloc((void**)&x, sizeof(float2) * n));
  CGM_HIP(hipMalloc((void**)&gq, sizeof(float2) * H * H));
  CGM_HIP(hipMalloc((void**)&chi, sizeof(float) * n));
  CGM_HIP(hipMemsetAsync(x, 0, sizeof(float2) * n, c->stream));
  CGM_HIP(hipMemsetAsync(gq, 0, sizeof(float2) * H * H, c->stream));
  CGM_HIP(hipMemsetAsync(chi, 0, sizeof(float) * n, c->stream));
  CGM_HIP(hipMemsetAsync(c->T32, 0, sizeof(float2) * (size_t)N * M, c->stream));
  hipEvent_t e0, e1;
  CGM_HIP(hipEventCreate(&e0));
  CGM_HIP(hipEventCreate(&e1));
  Plan P = c->plan;
  const int reps = 20;
  auto time_it = [&](auto&& launch) -> float {
    for (int r = 0; r < 3; ++r) launch();
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return ms / reps;
  };
  int rc = CGM_SUCCESS;
  const int cand[] = {1, 2, 4, 8};
  float best1 = 1e30f, best2 = 1e30f, best3 = 1e30f;
  int r1 = P.rows1, c2 = P.cols2, r3 = P.rows3;
  for (int k : cand) {
    if (k * nt > pass_max_threads(M, false) || (size_t)k * lds_len(M, false) * sizeof(float2) > 65536) continue;
    Plan Q = P;
    Q.rows1 = Q.rows3 = k;
    const float t1 = time_it([&] { rc |= launch_pass1<float>(Q, c->stream, x, nullptr, chi, nullptr, kModeForward, c->T32, c->tw32, nullptr); });
    const float t3 = time_it([&] { rc |= launch_pass3<float>(Q, c->stream, c->T32, x, chi, nullptr, c->vec_tmp ? c->vec_tmp : x, kModeForward, c->tw32, nullptr); });
    if (t1 < best1) { best1 = t1; r1 = k; }
    if (t3 < best3) { best3 = t3; r3 = k; }
    if (M % k == 0) {
      Q.cols2 = k;
      const float t2 = time_it([&] { rc |= launch_pass2<float, float2>(Q, c->stream, c->T32, gq, c->tw32, nullptr); });
      if (t2 < best2) { best2 = t2; c2 = k; }
    }
  }
  // M = 1024: the register-resident one-wave kernel of each pass against
  // the tuned LDS-staged radix-8 one, at batches of 1, 4 and 16
  // vectors; a pass takes the register-resident kernel from the smallest batch
  // at which it is faster at that batch and at every larger one
  int wide_from[3] = {0, 0, 0};
  float wide_us[3][3] = {}, lds_us[3][3] = {};
  if (!rc && M == 1024) {
    const int batches[3] = {1, 4, 16};
    const size_t nb = 16;
    float2 *Tb = nullptr, *xb = nullptr, *yb = nullptr;
    bool ok = hipMalloc((void**)&Tb, sizeof(float2) * (size_t)N * M * nb) == hipSuccess &&
              hipMalloc((void**)&xb, sizeof(float2) * (size_t)n * nb) == hipSuccess &&
              hipMalloc((void**)&yb, sizeof(float2) * (size_t)n * nb) == hipSuccess &&
              hipMemsetAsync(Tb, 0, sizeof(float2) * (size_t)N * M * nb, c->stream) == hipSuccess &&
              hipMemsetAsync(xb, 0, sizeof(float2) * (size_t)n * nb, c->stream) == hipSuccess;
    if (ok) {
      Plan Q = P;
      Q.rows1 = r1;
      Q.rows3 = r3;
      Q.cols2 = c2;
      BatchStrides bs;
      bs.x = n;
      bs.chi = 0;
      bs.y = n;
      bs.T = (long)N * M;
      for (int k = 0; k < 3; ++k) {
        const int B = batches[k];
        for (int wide = 0; wide < 2; ++wide) {
          Q.wide1_batch = Q.wide2_batch = Q.wide3_batch = wide;
          float(&dst)[3][3] = wide ? wide_us : lds_us;
          dst[0][k] = 1e3f * time_it([&] { rc |= launch_pass1<float>(Q, c->stream, xb, nullptr, chi, nullptr, kModeForward, Tb, c->tw32, nullptr, B, bs); });
          dst[1][k] = 1e3f * time_it([&] { rc |= launch_pass2<float, float2>(Q, c->stream, Tb, gq, c->tw32, nullptr, B, bs); });
          dst[2][k] = 1e3f * time_it([&] { rc |= launch_pass3<float>(Q, c->stream, Tb, xb, chi, nullptr, yb, kModeForward, c->tw32, nullptr, B, bs); });
        }
      }
      for (int q = 0; q < 3; ++q)
        for (int k = 2; k >= 0 && wide_us[q][k] < lds_us[q][k]; --k) wide_from[q] = batches[k];
    }
    (void)hipStreamSynchronize(c->stream);
    if (Tb) (void)hipFree(Tb);
    if (xb) (void)hipFree(xb);
    if (yb) (void)hipFree(yb);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(x);
  (void)hipFree(gq);
  (void)hipFree(chi);
  if (rc) return rc;
  c->plan.rows1 = r1;
  c->plan.rows3 = r3;
  c->plan.cols2 = c2;
  c->plan.wide1_batch = wide_from[0];
  c->plan.wide2_batch = wide_from[1];
  c->plan.wide3_batch = wide_from[2];
  for (int q = 0; q < 3; ++q)
    for (int k = 0; k < 3; ++k) {
      c->tuned_wide_us[q][k] = wide_us[q][k];
      c->tuned_lds_us[q][k] = lds_us[q][k];
    }
  c->tuned_us[0] = best1 * 1e3f;
  c->tuned_us[1] = best2 * 1e3f;
  c->tuned_us[2] = best3 * 1e3f;
  return CGM_SUCCESS;
}

extern "C" int cgm_get_tiles(cgm_ctx_t c, int* tiles, float* us) {
  CGM_CHECK(c && tiles, CGM_ERROR_INVALID_VALUE, "null argument");
  tiles[0] = c->plan.rows1;
  tiles[1] = c->plan.cols2;
  tiles[2] = c->plan.rows3;
  if (us)
    for (int k = 0; k < 3; ++k) us[k] = c->tuned_us[k];
  return CGM_SUCCESS;
}

extern "C" int cgm_set_radix_plan(cgm_ctx_t c, const int* wide_from_batch) {
  CGM_CHECK(c && wide_from_batch, CGM_ERROR_INVALID_VALUE, "null argument");
  for (int q = 0; q < 3; ++q)
    CGM_CHECK(wide_from_batch[q] >= 0, CGM_ERROR_INVALID_VALUE, "wide_from_batch[%d] < 0", q);
  const bool m1024 = c->M == 1024;
  c->plan.wide1_batch = m1024 ? wide_from_batch[0] : 0;
  c->plan.wide2_batch = m1024 ? wide_from_batch[1] : 0;
  c->plan.wide3_batch = m1024 ? wide_from_batch[2] : 0;
  return CGM_SUCCESS;
}

extern "C" int cgm_get_radix_plan(cgm_ctx_t c, int* wide_from_batch, float* wide_us, float* lds_us) {
  CGM_CHECK(c && wide_from_batch, CGM_ERROR_INVALID_VALUE, "null argument");
  wide_from_batch[0] = c->plan.wide1_batch;
  wide_from_batch[1] = c->plan.wide2_batch;
  wide_from_batch[2] = c->plan.wide3_batch;
  for (int q = 0; q < 3; ++q)
    for (int k = 0; k < 3; ++k) {
      if (wide_us) wide_us[3 * q + k] = c->tuned_wide_us[q][k];
      if (lds_us) lds_us[3 * q + k] = c->tuned_lds_us[q][k];
    }
  return CGM_SUCCESS;
}

extern "C" int cgm_set_stream(cgm_ctx_t c, void* stream) {
  CGM_CHECK(c, CGM_ERROR_INVALID_VALUE, "null context");
  c->stream = (hipStream_t)stream;
  return CGM_SUCCESS;
}

extern "C" int cgm_set_arnoldi_path(cgm_ctx_t c, int path) {
  CGM_CHECK(c && path >= 0 && path <= 3, CGM_ERROR_INVALID_VALUE, "path must be 0, 1, 2 or 3");
  c->path = path;
  return CGM_SUCCESS;
}

// Work-group shape of the fused kernels for B systems per launch: path 2
// forces 256-thread work-groups, path 3 1,024-thread ones.  Both shapes give
// a work-group 1024 / N rows, so a launch has B N^2 / 1024 work-groups;
// path 0 takes the 256-thread shape once that is at least 512 (two per CU,
// B N^2 >= 2^19), where its fuller waves pay (per system-step: N = 512,
// 4 systems 28.7 against 36.2 us; 16 systems at N = 192 / 256 / 512:
// 5.7 / 6.8 / 25.4 against 7.3 / 8.9 / 34.8), and the
// 1,024-thread shape below, whose 16 waves per work-group hide more latency
// when there are fewer work-groups (one system at N = 512: 54.6 against
// 70.4 us; 16 systems at N = 128: 3.05 against 3.46 us).
static int fused_small(const cgm_context* c, int B) {
  if (c->path == 2) return 1;
  if (c->path == 3) return 0;
  return (long)B * c->N * c->N >= (1L << 19) ? 1 : 0;
}

extern "C" int cgm_get_embedding(cgm_ctx_t c, int* M) {
  CGM_CHECK(c && M, CGM_ERROR_INVALID_VALUE, "null argument");
  *M = c->M;
  return CGM_SUCCESS;
}

extern "C" int cgm_spectrum_size(cgm_ctx_t c, size_t* n) {
  CGM_CHECK(c && n, CGM_ERROR_INVALID_VALUE, "null argument");
  const size_t H = (size_t)c->M / 2 + 1;
  *n = H * H;
  return CGM_SUCCESS;
}

extern "C" int cgm_synchronize(cgm_ctx_t c) {
  CGM_CHECK(c, CGM_ERROR_INVALID_VALUE, "null context");
  CGM_HIP(hipStreamSynchronize(c->stream));
  return CGM_SUCCESS;
}

// ---------------------------------------------------------------------------
// spectra
// ---------------------------------------------------------------------------
static int spectrum_common(cgm_context* c, const double2* blk, cgm_complex64* g32,
                           cgm_complex128* g64) {
  return spectrum_from_block(c->plan64, c->stream, blk, c->tw64, (float2*)g32, (double2*)g64);
}

extern "C" int cgm_spectrum_from_block(cgm_ctx_t c, const cgm_complex128* blk, cgm_complex64* g32,
                                       cgm_complex128* g64) {
  CGM_CHECK(c && blk && (g32 || g64), CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_HIP(hipSetDevice(c->device));
  int rc = spectrum_common(c, (const double2*)blk, g32, g64);
  if (rc) return rc;
  CGM_HIP(hipStreamSynchronize(c->stream));
  return CGM_SUCCESS;
}

extern "C" int cgm_spectrum_from_gprep(cgm_ctx_t c, const cgm_complex64* gprep, cgm_complex64* g32) {
  CGM_CHECK(c && gprep && g32, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_HIP(hipSetDevice(c->device));
  const int N = c->N, Mp = 2 * N - 1;
  // symmetry check on the host copy (one-off setup)
  std::vector<cgm_complex64> h((size_t)Mp * Mp);
  CGM_HIP(hipMemcpy(h.data(), gprep, sizeof(cgm_complex64) * h.size(), hipMemcpyDeviceToHost));
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      const cgm_complex64 a = h[(size_t)i * Mp + j];
      const size_t mir[3] = {(size_t)((Mp - i) % Mp) * Mp + j, (size_t)i * Mp + (Mp - j) % Mp,
                             (size_t)((Mp - i) % Mp) * Mp + (Mp - j) % Mp};
      for (size_t q : mir) {
        const cgm_complex64 b = h[q];
        if (b.re != a.re || b.im != a.im) {
          cgm_set_last_error("cgm_spectrum_from_gprep: G_prep not even at (%d,%d)", i, j);
          return CGM_ERROR_NOT_SYMMETRIC;
        }
      }
    }
  double2* blk = nullptr;
  CGM_HIP(hipMalloc((void**)&blk, sizeof(double2) * (size_t)N * N));
  int rc = gprep_block(c->stream, (const float2*)gprep, blk, N);
  if (!rc) rc = spectrum_common(c, blk, g32, nullptr);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(blk);
  return rc;
}

static int spectrum_generated(cgm_context* c, int kind, double k, double sr, double sc,
                              cgm_complex64* g32, cgm_complex128* g64) {
  CGM_CHECK(c && (g32 || g64), CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(k > 0 && isfinite(k), CGM_ERROR_INVALID_VALUE, "wave number must be positive");
  CGM_HIP(hipSetDevice(c->device));
  const int N = c->N;
  double2* blk = nullptr;
  CGM_HIP(hipMalloc((void**)&blk, sizeof(double2) * (size_t)N * N));
  int rc = green_block(c->stream, blk, N, kind, k, sr, sc);
  if (!rc) rc = spectrum_common(c, blk, g32, g64);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(blk);
  return rc;
}

extern "C" int cgm_spectrum_reference(cgm_ctx_t c, double k, double sr, double sc,
                                      cgm_complex64* g32, cgm_complex128* g64) {
  // the shifted distance must stay positive at every offset (H0 is singular at 0)
  CGM_CHECK(isfinite(sr) && isfinite(sc) && (fabs(sr - rint(sr)) > 0 || fabs(sc - rint(sc)) > 0),
            CGM_ERROR_INVALID_VALUE, "cgm_spectrum_reference: shifts must not both be integers");
  return spectrum_generated(c, 1, k, sr, sc, g32, g64);
}

extern "C" int cgm_spectrum_helmholtz(cgm_ctx_t c, double kh, cgm_complex64* g32,
                                      cgm_complex128* g64) {
  return spectrum_generated(c, 0, kh, 0.0, 0.0, g32, g64);
}

static int block_generated(cgm_context* c, int kind, double k, double sr, double sc, cgm_complex128* blk) {
  CGM_CHECK(c && blk, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(k > 0 && isfinite(k), CGM_ERROR_INVALID_VALUE, "wave number must be positive");
  CGM_HIP(hipSetDevice(c->device));
  return green_block(c->stream, (double2*)blk, c->N, kind, k, sr, sc);
}

extern "C" int cgm_block_helmholtz(cgm_ctx_t c, double kh, cgm_complex128* blk) {
  return block_generated(c, 0, kh, 0.0, 0.0, blk);
}

extern "C" int cgm_block_reference(cgm_ctx_t c, double k, double sr, double sc, cgm_complex128* blk) {
  return block_generated(c, 1, k, sr, sc, blk);
}

// ---------------------------------------------------------------------------
// operator
// ---------------------------------------------------------------------------
static int matvec32(cgm_context* c, const float2* gq, const float* chi, const float2* x,
                    const float2* b, float2* y, int mode, const double* norm_ptr, float2* xw,
                    const int* done) {
  int rc = launch_pass1<float>(c->plan, c->stream, x, xw, chi, norm_ptr, mode, c->T32, c->tw32, done);
  if (!rc) rc = launch_pass2<float, float2>(c->plan, c->stream, c->T32, gq, c->tw32, done);
  // pass 3 reads the (possibly rescaled) input written back by pass 1
  const float2* xin = xw ? xw : x;
  if (!rc) rc = launch_pass3<float>(c->plan, c->stream, c->T32, xin, chi, b, y, mode, c->tw32, done);
  return rc;
}

static int ensure_T64(cgm_context* c) {
  if (!c->T64) CGM_HIP(hipMalloc((void**)&c->T64, sizeof(double2) * (size_t)c->N * c->M));
  return CGM_SUCCESS;
}

static int matvec64(cgm_context* c, const double2* gq, const double* chi, const double2* x,
                    const double2* b, double2* y, int mode) {
  int rc = ensure_T64(c);
  if (!rc) rc = launch_pass1<double>(c->plan64, c->stream, x, nullptr, chi, nullptr, mode, c->T64, c->tw64, nullptr);
  if (!rc) rc = launch_pass2<double, double2>(c->plan64, c->stream, c->T64, gq, c->tw64, nullptr);
  if (!rc) rc = launch_pass3<double>(c->plan64, c->stream, c->T64, x, chi, b, y, mode, c->tw64, nullptr);
  return rc;
}

extern "C" int cgm_matvec(cgm_ctx_t c, const cgm_complex64* gamma, const float* chi,
                          const cgm_complex64* x, cgm_complex64* y, int mode) {
  CGM_CHECK(c && gamma && x && y, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(mode >= 0 && mode <= 2, CGM_ERROR_INVALID_VALUE, "mode %d", mode);
  CGM_CHECK(mode == CGM_OP_GREEN || chi, CGM_ERROR_INVALID_VALUE, "null contrast");
  CGM_HIP(hipSetDevice(c->device));
  return matvec32(c, (const float2*)gamma, chi, (const float2*)x, nullptr, (float2*)y, mode,
                  nullptr, nullptr, nullptr);
}

static int ensure_T32B(cgm_context* c, int B) {
  if (c->T32B_n >= B) return CGM_SUCCESS;
  if (c->T32B) CGM_HIP(hipFree(c->T32B));
  c->T32B = nullptr;
  c->T32B_n = 0;
  CGM_HIP(hipMalloc((void**)&c->T32B, sizeof(float2) * (size_t)c->N * c->M * B));
  c->T32B_n = B;
  return CGM_SUCCESS;
}

extern "C" int cgm_matvec_batched(cgm_ctx_t c, const cgm_complex64* gamma, const float* chi, long chi_stride,
                                  const cgm_complex64* x, cgm_complex64* y, int batch, int mode) {
  CGM_CHECK(c && gamma && x && y, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(batch >= 1, CGM_ERROR_INVALID_VALUE, "batch %d < 1", batch);
  CGM_CHECK(mode >= 0 && mode <= 2, CGM_ERROR_INVALID_VALUE, "mode %d", mode);
  CGM_CHECK(mode == CGM_OP_GREEN || chi, CGM_ERROR_INVALID_VALUE, "null contrast");
  CGM_CHECK(chi_stride == 0 || chi_stride >= (long)c->N * c->N, CGM_ERROR_INVALID_VALUE,
            "chi_stride must be 0 (shared map) or >= N*N");
  CGM_HIP(hipSetDevice(c->device));
  int rc = ensure_T32B(c, batch);
  if (rc) return rc;
  const long n = (long)c->N * c->N;
  BatchStrides bs;
  bs.x = n;
  bs.chi = chi_stride;
  bs.y = n;
  bs.T = (long)c->N * c->M;
  rc = launch_pass1<float>(c->plan, c->stream, (const float2*)x, nullptr, chi, nullptr, mode, c->T32B, c->tw32,
                           nullptr, batch, bs);
  if (!rc) rc = launch_pass2<float, float2>(c->plan, c->stream, c->T32B, (const float2*)gamma, c->tw32, nullptr,
                                            batch, bs);
  if (!rc) rc = launch_pass3<float>(c->plan, c->stream, c->T32B, (const float2*)x, chi, nullptr, (float2*)y, mode,
                                    c->tw32, nullptr, batch, bs);
  return rc;
}

extern "C" int cgm_matvec_c128(cgm_ctx_t c, const cgm_complex128* gamma, const double* chi,
                               const cgm_complex128* x, cgm_complex128* y, int mode) {
  CGM_CHECK(c && gamma && x && y, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(mode >= 0 && mode <= 2, CGM_ERROR_INVALID_VALUE, "mode %d", mode);
  CGM_CHECK(mode == CGM_OP_GREEN || chi, CGM_ERROR_INVALID_VALUE, "null contrast");
  CGM_HIP(hipSetDevice(c->device));
  return matvec64(c, (const double2*)gamma, chi, (const double2*)x, nullptr, (double2*)y, mode);
}

// ---------------------------------------------------------------------------
// GMRES
// ---------------------------------------------------------------------------
static int read_state(cgm_context* c) {
  CGM_HIP(hipMemcpyAsync(c->hst, c->st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
  CGM_HIP(hipStreamSynchronize(c->stream));
  return CGM_SUCCESS;
}

// Core solver on float2 device vectors.  hist: optional device array,
// hist[0] = beta0, hist[t] = estimate after t steps.
// One restart cycle of plain GMRES from the residual in V_0 (its norm in
// st->beta), up to the Givens-rotated least-squares problem; the caller
// applies x += V y.  Fused lagged-CGS2 steps where supported.
static int plain_cycle(cgm_context* c, const float2* gq, const float* chi, int fwd, int m, double* hist,
                       int total) {
  KrylovWs& ws = c->ws;
  const long n = ws.n;
  const bool fused = c->path != 1 && fused_supported(c->N, m);
  int rc;
  if (fused) {
    // r = b - A x is in V_0; u_0 = r, first matvec's passes 1-2
    float2* U0 = ws.U;
    CGM_HIP(hipMemcpyAsync(U0, ws.V, sizeof(float2) * n, hipMemcpyDeviceToDevice, c->stream));
    rc = launch_pass1<float>(c->plan, c->stream, U0, nullptr, chi, nullptr, fwd, c->T32, c->tw32, &c->st->done);
    if (!rc) rc = launch_pass2<float, float2>(c->plan, c->stream, c->T32, gq, c->tw32, &c->st->done);
    if (rc) return rc;
    FusedLaunch L{};
    L.N = c->N; L.m = m; L.ld = ws.m; L.mode = fwd; L.chi = chi; L.V = ws.V; L.Z = ws.Z;
    L.T = c->T32; L.tw = c->tw32; L.part = ws.part; L.H = ws.H; L.Rm = ws.R; L.cs = ws.cs;
    L.sn = ws.sn; L.g = ws.g; L.h1p = ws.h1p; L.a1 = ws.a1; L.a2 = ws.a2; L.scal = ws.scal;
    L.st = c->st; L.hist = hist ? hist + total : nullptr;
    L.small = fused_small(c, 1);
    for (int j = 0; j <= m; ++j) {
      L.j = j;
      L.Uj = ws.U + (long)(j & 1) * n;
      L.Unext = ws.U + (long)((j + 1) & 1) * n;
      if ((rc = launch_fused(c->stream, 0, L))) return rc;
      if (j == m) break;
      if ((rc = launch_fused(c->stream, 1, L))) return rc;
      if (j + 1 == m) continue;  // u_m needs no matvec
      if ((rc = launch_pass2<float, float2>(c->plan, c->stream, c->T32, gq, c->tw32, &c->st->done))) return rc;
    }
  } else {
    for (int i = 0; i < m; ++i) {
      float2* vi = ws.V + (long)i * n;
      const double* nrm = (i == 0) ? &c->st->beta : &c->st->hn;
      rc = matvec32(c, gq, chi, vi, nullptr, vi + n, fwd, nrm, vi, &c->st->done);
      if (rc) return rc;
      rc = launch_cgs2_step(c->stream, ws, i, c->st, hist ? hist + 1 + total : nullptr);
      if (rc) return rc;
    }
  }
  return CGM_SUCCESS;
}

static int gmres_core(cgm_context* c, const float2* gq, const float* chi, const float2* b,
                      float2* x, const cgm_gmres_params* p, cgm_gmres_result* res, double* hist) {
  const long n = (long)c->N * c->N;
  const int m = p->restart;
  int rc = krylov_alloc(c->ws, n, m, c->stream);
  if (rc) return rc;
  KrylovWs& ws = c->ws;
  const int fwd = p->adjoint ? kModeAdjoint : kModeForward;
  const int resmode = p->adjoint ? kModeResidualAdj : kModeResidual;
  // ||b|| and the absolute tolerance
  CGM_HIP(hipMemsetAsync(&c->st->status, 0, sizeof(int), c->stream));
  if ((rc = launch_norm_init(c->stream, ws, b, c->st, 1, p->tol))) return rc;
  if (p->x0_zero) CGM_HIP(hipMemsetAsync(x, 0, sizeof(float2) * n, c->stream));
  int total = 0, cycles = 0;
  res->converged = 0;
  for (;;) {
    // true residual r = b - A x into V_0
    if (p->x0_zero && cycles == 0) {
      CGM_HIP(hipMemcpyAsync(ws.V, b, sizeof(float2) * n, hipMemcpyDeviceToDevice, c->stream));
    } else {
      rc = matvec32(c, gq, chi, x, b, ws.V, resmode, nullptr, nullptr, nullptr);
      if (rc) return rc;
    }
    if ((rc = launch_norm_init(c->stream, ws, ws.V, c->st, 0, 0.0))) return rc;
    if ((rc = read_state(c))) return rc;
    const DevState& h = *c->hst;
    if (cycles == 0) {
      res->beta0 = h.beta;
      res->bnorm = h.bnorm;
      if (hist) CGM_HIP(hipMemcpyAsync(hist, &h.beta, sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    res->rel_residual = h.bnorm > 0 ? h.beta / h.bnorm : 0.0;
    if (h.status) {
      cgm_set_last_error("cgm_gmres: non-finite residual");
      res->iterations = total;
      res->cycles = cycles;
      return CGM_ERROR_NONFINITE;
    }
    if (h.bnorm == 0.0 || h.beta <= p->tol * h.bnorm) { res->converged = 1; break; }
    if (cycles >= p->max_cycles) break;
    // one restart cycle, enqueued without host synchronisation
    if ((rc = plain_cycle(c, gq, chi, fwd, m, hist, total))) return rc;
    if ((rc = launch_backsub_update(c->stream, ws, c->st, x))) return rc;
    if ((rc = read_state(c))) return rc;
    total += c->hst->k;
    ++cycles;
    if (c->hst->status) {
      cgm_set_last_error("cgm_gmres: non-finite value in the Arnoldi process");
      res->iterations = total;
      res->cycles = cycles;
      return CGM_ERROR_NONFINITE;
    }
  }
  res->iterations = total;
  res->cycles = cycles;
  return res->converged ? CGM_SUCCESS : CGM_ERROR_NOT_CONVERGED;
}

// Exactly one cycle of m steps from x, entered only when ||b - A x|| >
// tolerance (absolute) and never left early: the control of the reference's
// unrestarted solver (B/CUDA_GMRES/GMRES.cuh:153,253).  m + 1 matvecs (the
// initial residual and one per step); hist[0] = ||r0||, hist[t] = estimate
// after t steps.  res->iterations = steps done (0 when not entered).
int cgm::gmres_fixed_cycle(cgm_context* c, const float2* gq, const float* chi, const float2* b, float2* x,
                           int m, int adjoint, double tolerance, double* hist, cgm_gmres_result* res) {
  const long n = (long)c->N * c->N;
  int rc = krylov_alloc(c->ws, n, m, c->stream);
  if (rc) return rc;
  KrylovWs& ws = c->ws;
  memset(res, 0, sizeof(*res));
  CGM_HIP(hipMemsetAsync(&c->st->status, 0, sizeof(int), c->stream));
  // ||b||, and a zero absolute tolerance inside the cycle
  if ((rc = launch_norm_init(c->stream, ws, b, c->st, 1, 0.0))) return rc;
  rc = matvec32(c, gq, chi, x, b, ws.V, adjoint ? kModeResidualAdj : kModeResidual, nullptr, nullptr, nullptr);
  if (rc) return rc;
  if ((rc = launch_norm_init(c->stream, ws, ws.V, c->st, 0, 0.0))) return rc;
  if ((rc = read_state(c))) return rc;
  res->beta0 = c->hst->beta;
  res->bnorm = c->hst->bnorm;
  res->rel_residual = res->bnorm > 0 ? res->beta0 / res->bnorm : 0.0;
  if (hist) CGM_HIP(hipMemcpyAsync(hist, &c->hst->beta, sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (c->hst->status) {
    cgm_set_last_error("non-finite initial residual");
    return CGM_ERROR_NONFINITE;
  }
  if (!(res->beta0 > tolerance)) return CGM_SUCCESS;
  if ((rc = plain_cycle(c, gq, chi, adjoint ? kModeAdjoint : kModeForward, m, hist, 0))) return rc;
  if ((rc = launch_backsub_update(c->stream, ws, c->st, x))) return rc;
  if ((rc = read_state(c))) return rc;
  res->iterations = c->hst->k;
  res->cycles = 1;
  res->rel_residual = res->bnorm > 0 ? c->hst->res_est / res->bnorm : 0.0;
  if (c->hst->status) {
    cgm_set_last_error("non-finite value in the Arnoldi process");
    return CGM_ERROR_NONFINITE;
  }
  return CGM_SUCCESS;
}

static int check_params(const cgm_gmres_params* p) {
  CGM_CHECK(p, CGM_ERROR_INVALID_VALUE, "null params");
  CGM_CHECK(p->restart >= 1 && p->restart <= kMaxRestart, CGM_ERROR_INVALID_VALUE,
            "restart %d outside [1, %d]", p->restart, kMaxRestart);
  CGM_CHECK(p->max_cycles >= 0, CGM_ERROR_INVALID_VALUE, "max_cycles < 0");
  CGM_CHECK(p->tol >= 0 && isfinite(p->tol), CGM_ERROR_INVALID_VALUE, "bad tolerance");
  return CGM_SUCCESS;
}

extern "C" int cgm_gmres(cgm_ctx_t c, const cgm_complex64* gamma, const float* chi,
                         const cgm_complex64* b, cgm_complex64* x, const cgm_gmres_params* p,
                         cgm_gmres_result* res, double* hist) {
  CGM_CHECK(c && gamma && chi && b && x && res, CGM_ERROR_INVALID_VALUE, "null argument");
  int rc = check_params(p);
  if (rc) return rc;
  CGM_HIP(hipSetDevice(c->device));
  memset(res, 0, sizeof(*res));
  return gmres_core(c, (const float2*)gamma, chi, (const float2*)b, (float2*)x, p, res, hist);
}

// ---------------------------------------------------------------------------
// Batched GMRES: B systems of one grid size in lockstep
// ---------------------------------------------------------------------------
__global__ void k_copy_strided(const float2* __restrict__ src, long ss, float2* __restrict__ dst, long ds, long n) {
  const long k = blockIdx.y;
  src += k * ss;
  dst += k * ds;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

static int copy_strided(hipStream_t st, const float2* src, long ss, float2* dst, long ds, long n, int B) {
  const long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_copy_strided, dim3((unsigned)(blocks < 1024 ? blocks : 1024), B), dim3(256), 0, st, src, ss,
                     dst, ds, n);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

static int read_states(cgm_context* c, int B, hipStream_t st) {
  CGM_HIP(hipMemcpyAsync(c->hstB, c->stB, sizeof(DevState) * B, hipMemcpyDeviceToHost, st));
  CGM_HIP(hipStreamSynchronize(st));
  return CGM_SUCCESS;
}

extern "C" int cgm_set_batched_graphs(cgm_ctx_t c, int enable) {
  CGM_CHECK(c, CGM_ERROR_INVALID_VALUE, "null context");
  c->use_graphs = enable ? 1 : 0;
  return CGM_SUCCESS;
}

extern "C" int cgm_gmres_batched(cgm_ctx_t c, const cgm_complex64* gamma, const float* chi, long chi_stride,
                                 const cgm_complex64* b, cgm_complex64* x, int batch, const cgm_gmres_params* p,
                                 cgm_gmres_result* results, double* hist) {
  CGM_CHECK(c && gamma && chi && b && x && results, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(batch >= 1, CGM_ERROR_INVALID_VALUE, "batch %d < 1", batch);
  CGM_CHECK(chi_stride == 0 || chi_stride >= (long)c->N * c->N, CGM_ERROR_INVALID_VALUE,
            "chi_stride must be 0 (shared map) or >= N*N");
  int rc = check_params(p);
  if (rc) return rc;
  CGM_CHECK(!p->adjoint, CGM_ERROR_UNSUPPORTED, "batched adjoint solves are not supported");
  CGM_HIP(hipSetDevice(c->device));
  memset(results, 0, sizeof(cgm_gmres_result) * batch);
  const long n = (long)c->N * c->N;
  const int m = p->restart;
  const int B = batch;
  const long hstride = (long)m * (p->max_cycles > 0 ? p->max_cycles : 1) + 1;
  if (!fused_supported(c->N, m)) {
    // grid sizes without the fused kernels: the systems one after another
    for (int k = 0; k < B && !rc; ++k) {
      const int r = gmres_core(c, (const float2*)gamma, chi + k * chi_stride, (const float2*)b + k * n,
                               (float2*)x + k * n, p, &results[k], hist ? hist + k * hstride : nullptr);
      if (r && r != CGM_ERROR_NOT_CONVERGED) rc = r;
    }
    if (rc) return rc;
    for (int k = 0; k < B; ++k)
      if (!results[k].converged) return CGM_ERROR_NOT_CONVERGED;
    return CGM_SUCCESS;
  }
  if ((rc = krylov_alloc(c->ws, n, m, c->stream, B))) return rc;
  KrylovWs& ws = c->ws;
  if (c->stB_n < B) {
    if (c->stB) CGM_HIP(hipFree(c->stB));
    if (c->hstB) CGM_HIP(hipHostFree(c->hstB));
    c->stB = nullptr;
    c->hstB = nullptr;
    c->stB_n = 0;
    CGM_HIP(hipMalloc((void**)&c->stB, sizeof(DevState) * B));
    CGM_HIP(hipHostMalloc((void**)&c->hstB, sizeof(DevState) * B));
    c->stB_n = B;
  }
  const long NM = (long)c->N * c->M;
  if ((rc = ensure_T32B(c, B))) return rc;
  const long sst = (long)sizeof(DevState);
  const float2* gq = (const float2*)gamma;
  const float2* bb = (const float2*)b;
  float2* xx = (float2*)x;
  DevState* stB = c->stB;
  CGM_HIP(hipMemsetAsync(stB, 0, sizeof(DevState) * B, c->stream));
  if (hist) CGM_HIP(hipMemsetAsync(hist, 0, sizeof(double) * hstride * B, c->stream));
  if ((rc = launch_norm_init(c->stream, ws, bb, stB, 1, p->tol, n, sst, B))) return rc;
  if (p->x0_zero) CGM_HIP(hipMemsetAsync(xx, 0, sizeof(float2) * n * B, c->stream));
  // residual r = b - A x of every system into its V_0
  BatchStrides rs;
  rs.x = n;
  rs.chi = chi_stride;
  rs.b = n;
  rs.y = ws.sBasis;
  rs.T = NM;
  rs.st = sst;
  // u_0 = r and the first matvec's passes 1-2
  BatchStrides us = rs;
  us.x = 2 * n;
  // one restart cycle of every live system (converged ones are skipped by
  // their done flags), then the next cycle's residual and its norm; the
  // launch sequence and its pointers are the same every cycle
  auto enqueue_cycle = [&](hipStream_t s, int cyc) -> int {
    int r = copy_strided(s, ws.V, ws.sBasis, ws.U, 2 * n, n, B);
    if (!r) r = launch_pass1<float>(c->plan, s, ws.U, nullptr, chi, nullptr, kModeForward, c->T32B, c->tw32,
                                    &stB->done, B, us);
    if (!r) r = launch_pass2<float, float2>(c->plan, s, c->T32B, gq, c->tw32, &stB->done, B, us);
    if (r) return r;
    FusedLaunch L{};
    L.N = c->N; L.m = m; L.ld = ws.m; L.mode = kModeForward; L.chi = chi; L.V = ws.V; L.Z = ws.Z;
    L.T = c->T32B; L.tw = c->tw32; L.part = ws.part; L.H = ws.H; L.Rm = ws.R; L.cs = ws.cs;
    L.sn = ws.sn; L.g = ws.g; L.h1p = ws.h1p; L.a1 = ws.a1; L.a2 = ws.a2; L.scal = ws.scal;
    L.st = stB; L.hist = hist ? hist + (long)cyc * m : nullptr;
    L.B = B; L.sChi = chi_stride; L.sBasis = ws.sBasis; L.sU = 2 * n; L.sVec = n; L.sT = NM;
    L.sPart = ws.sPart; L.sSmall = ws.sSmall; L.sSt = sst; L.sHist = hstride;
    L.small = fused_small(c, B);
    for (int j = 0; j <= m; ++j) {
      L.j = j;
      L.Uj = ws.U + (long)(j & 1) * n;
      L.Unext = ws.U + (long)((j + 1) & 1) * n;
      if ((r = launch_fused(s, 0, L))) return r;
      if (j == m) break;
      if ((r = launch_fused(s, 1, L))) return r;
      if (j + 1 == m) continue;  // u_m needs no matvec
      if ((r = launch_pass2<float, float2>(c->plan, s, c->T32B, gq, c->tw32, &stB->done, B, us))) return r;
    }
    if ((r = launch_backsub_update(s, ws, stB, xx, nullptr, n, sst, B))) return r;
    r = launch_pass1<float>(c->plan, s, xx, nullptr, chi, nullptr, kModeResidual, c->T32B, c->tw32, nullptr, B, rs);
    if (!r) r = launch_pass2<float, float2>(c->plan, s, c->T32B, gq, c->tw32, nullptr, B, rs);
    if (!r) r = launch_pass3<float>(c->plan, s, c->T32B, xx, chi, bb, ws.V, kModeResidual, c->tw32, nullptr, B, rs);
    if (!r) r = launch_norm_init(s, ws, ws.V, stB, 0, 0.0, ws.sBasis, sst, B);
    return r;
  };
  // first residual, eagerly on the context stream
  if (p->x0_zero) {
    if ((rc = copy_strided(c->stream, bb, n, ws.V, ws.sBasis, n, B))) return rc;
  } else {
    rc = launch_pass1<float>(c->plan, c->stream, xx, nullptr, chi, nullptr, kModeResidual, c->T32B, c->tw32,
                             nullptr, B, rs);
    if (!rc) rc = launch_pass2<float, float2>(c->plan, c->stream, c->T32B, gq, c->tw32, nullptr, B, rs);
    if (!rc)
      rc = launch_pass3<float>(c->plan, c->stream, c->T32B, xx, chi, bb, ws.V, kModeResidual, c->tw32, nullptr, B,
                               rs);
    if (rc) return rc;
  }
  if ((rc = launch_norm_init(c->stream, ws, ws.V, stB, 0, 0.0, ws.sBasis, sst, B))) return rc;
  if ((rc = read_states(c, B, c->stream))) return rc;
  // captured cycles run on the context's own non-blocking stream (a capture
  // needs one; the caller's stream may be the null stream)
  const bool graph = c->use_graphs && !hist;
  hipStream_t cs = c->stream;
  hipGraphExec_t exec = nullptr;
  if (graph) {
    if (!c->gstream) CGM_HIP(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
    cs = c->gstream;
  }
  std::vector<int> done(B, 0);
  int cycles = 0;
  for (;;) {
    int active = 0;
    for (int k = 0; k < B; ++k) {
      const DevState& h = c->hstB[k];
      cgm_gmres_result& r = results[k];
      if (cycles == 0) {
        r.beta0 = h.beta;
        r.bnorm = h.bnorm;
        if (hist) CGM_HIP(hipMemcpyAsync(hist + k * hstride, &r.beta0, sizeof(double), hipMemcpyHostToDevice,
                                         c->stream));
      } else if (!done[k]) {
        r.iterations += h.k_last;
        r.cycles += 1;
      }
      if (done[k]) continue;
      r.rel_residual = h.bnorm > 0 ? h.beta / h.bnorm : 0.0;
      if (h.status) {
        done[k] = 1;
        continue;
      }
      if (h.bnorm == 0.0 || h.beta <= p->tol * h.bnorm) {
        r.converged = 1;
        done[k] = 1;
        continue;
      }
      ++active;
    }
    if (!active || cycles >= p->max_cycles) break;
    if (graph) {
      if (!exec) {
        hipGraph_t gr = nullptr;
        CGM_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        rc = enqueue_cycle(cs, 0);
        const hipError_t ec = hipStreamEndCapture(cs, &gr);
        if (rc) {
          if (gr) (void)hipGraphDestroy(gr);
          return rc;
        }
        CGM_HIP(ec);
        const hipError_t ei = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        CGM_HIP(ei);
      }
      const hipError_t el = hipGraphLaunch(exec, cs);
      if (el != hipSuccess) {
        (void)hipGraphExecDestroy(exec);
        CGM_HIP(el);
      }
    } else if ((rc = enqueue_cycle(cs, cycles))) {
      return rc;
    }
    rc = read_states(c, B, cs);
    if (rc) {
      if (exec) (void)hipGraphExecDestroy(exec);
      return rc;
    }
    ++cycles;
  }
  if (exec) CGM_HIP(hipGraphExecDestroy(exec));
  int any_nonfinite = 0, all_conv = 1;
  for (int k = 0; k < B; ++k) {
    if (c->hstB[k].status) any_nonfinite = 1;
    if (!results[k].converged) all_conv = 0;
  }
  if (any_nonfinite) {
    cgm_set_last_error("cgm_gmres_batched: non-finite value in at least one system");
    return CGM_ERROR_NONFINITE;
  }
  return all_conv ? CGM_SUCCESS : CGM_ERROR_NOT_CONVERGED;
}

// ---------------------------------------------------------------------------
// Flexible GMRES with a right preconditioner
// ---------------------------------------------------------------------------
// Identity preconditioner: v_i /= *nrm in place and z_i = v_i.
__global__ void k_norm_copy(float2* __restrict__ v, const double* __restrict__ nrm, float2* __restrict__ z, long n,
                            const int* done) {
  if (*done) return;
  const double d = *nrm;
  const float r = d > 0.0 ? (float)(1.0 / d) : 0.f;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    float2 a = v[e];
    a.x *= r;
    a.y *= r;
    v[e] = a;
    z[e] = a;
  }
}

// Identity preconditioner of the flexible fused steps: z = v / *nrm (v kept).
__global__ void k_scale_copy(const float2* __restrict__ v, const double* __restrict__ nrm, float2* __restrict__ z,
                             long n, const int* done) {
  if (*done) return;
  const double d = *nrm;
  const float r = d > 0.0 ? (float)(1.0 / d) : 0.f;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float2 a = v[e];
    z[e] = make_float2(a.x * r, a.y * r);
  }
}

// The preconditioned basis Z, at least len elements.
static int ensure_zb(cgm_context* c, long len) {
  if (c->zb && c->zb_len >= len) return CGM_SUCCESS;
  if (c->zb) CGM_HIP(hipFree(c->zb));
  c->zb = nullptr;
  c->zb_len = 0;
  CGM_HIP(hipMalloc((void**)&c->zb, sizeof(float2) * (size_t)len));
  c->zb_len = len;
  return CGM_SUCCESS;
}

// z /= *nrm for system blockIdx.y (z + y * sz, nrm + y * sn bytes): the
// identity preconditioner of batched flexible steps.
__global__ void k_scale_inplace(float2* __restrict__ z, long sz, const double* nrm, long sn, long n) {
  z += (long)blockIdx.y * sz;
  const double d = *boff(nrm, (long)blockIdx.y * sn);
  const float r = d > 0.0 ? (float)(1.0 / d) : 0.f;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float2 a = z[e];
    z[e] = make_float2(a.x * r, a.y * r);
  }
}

// z += rho z2 per system, rho = st->norm_out of that system (the two-level
// preconditioner's correction).
__global__ void k_add_scaled(float2* __restrict__ z, long sz, const float2* __restrict__ z2, long n,
                             const DevState* st, long sst) {
  z += (long)blockIdx.y * sz;
  z2 += (long)blockIdx.y * n;
  st = boff(st, (long)blockIdx.y * sst);
  if (st->done) return;
  const float rho = (float)st->norm_out;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float2 a = z[e], b = z2[e];
    z[e] = make_float2(a.x + rho * b.x, a.y + rho * b.y);
  }
}

// Scratch of the two-level preconditioner, `systems` vectors each.
static int ensure_pc(cgm_context* c, long systems) {
  const long len = systems * (long)c->N * c->N;
  if (c->pc_v && c->pc_len >= len) return CGM_SUCCESS;
  void* ptrs[] = {c->pc_v, c->pc_r, c->pc_z};
  for (void* q : ptrs)
    if (q) CGM_HIP(hipFree(q));
  c->pc_v = c->pc_r = c->pc_z = nullptr;
  c->pc_len = 0;
  CGM_HIP(hipMalloc((void**)&c->pc_v, sizeof(float2) * 2 * (size_t)len));
  CGM_HIP(hipMalloc((void**)&c->pc_r, sizeof(float2) * (size_t)len));
  CGM_HIP(hipMalloc((void**)&c->pc_z, sizeof(float2) * (size_t)len));
  c->pc_len = len;
  return CGM_SUCCESS;
}

// z = P(v / nrm) for B systems with the strides of io; system k's state is
// dst + k sst bytes (its done flag skips the work once set).
//   tiles only:   z = T v'                              (v' = v / nrm)
//   network only: z = F(v')
//   both (two-level, multiplicative): z1 = T v', r = v' - A z1,
//                 z = z1 + ||r|| F(r / ||r||)
// The network is thus a global correction on what the tile inverses leave:
// it is trained on the residuals r of the local solves.  The two-level form
// needs ensure_pc(c, B) and uses the context's B-system FFT scratch `Tm`.
static int precond_apply(cgm_context* c, const float2* gq, const Precond& P, hipStream_t st, const FnoIO& io, int B,
                         DevState* dst, long sst, float2* Tm) {
  const int* done = &dst->done;
  if (!P.tiles) return fno_forward_io(P.net, st, io, B, false, B == 1 ? done : nullptr);
  if (!P.net) return tiles_apply_io(P.tiles, st, io, B, done, sst);
  const long n = (long)c->N * c->N;
  // v' (normalised input): the caller's write-back target, or the scratch
  // (the write-back uses the input's stride io.sV; the scratch holds
  // 2 n per system, the largest input stride of the flexible solvers)
  FnoIO t = io;
  const float2* vh = io.v;
  const long svh = io.sV;
  if (io.nrm) {
    if (!io.vw) {
      if ((long)(B - 1) * io.sV + n > 2 * c->pc_len) {
        cgm_set_last_error("two-level preconditioner: input stride %ld exceeds the scratch", io.sV);
        return CGM_ERROR_UNSUPPORTED;
      }
      t.vw = c->pc_v;
    }
    vh = t.vw;
  }
  int rc = tiles_apply_io(P.tiles, st, t, B, done, sst);
  if (rc) return rc;
  // r = v' - A z1 into pc_r
  BatchStrides bs;
  bs.x = io.sZ;
  bs.chi = io.sChi;
  bs.b = svh;
  bs.y = n;
  bs.T = (long)c->N * c->M;
  bs.st = sst;
  rc = launch_pass1<float>(c->plan, st, io.z, nullptr, io.chi, nullptr, kModeResidual, Tm, c->tw32, done, B, bs);
  if (!rc) rc = launch_pass2<float, float2>(c->plan, st, Tm, gq, c->tw32, done, B, bs);
  if (!rc)
    rc = launch_pass3<float>(c->plan, st, Tm, io.z, io.chi, vh, c->pc_r, kModeResidual, c->tw32, done, B, bs);
  // ||r|| into norm_out of each system
  if (!rc) rc = launch_norm_init(st, c->ws, c->pc_r, dst, 2, 0.0, n, sst, B);
  if (rc) return rc;
  FnoIO f;
  f.chi = io.chi;
  f.sChi = io.sChi;
  f.v = c->pc_r;
  f.sV = n;
  f.nrm = &dst->norm_out;
  f.sNrm = sst;
  f.z = c->pc_z;
  f.sZ = n;
  if ((rc = fno_forward_io(P.net, st, f, B, false, B == 1 ? done : nullptr))) return rc;
  hipLaunchKernelGGL(k_add_scaled, dim3(c->ws.G, B), dim3(kKB), 0, st, io.z, io.sZ, c->pc_z, n, dst, sst);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

static FnoIO single_io(long n, const float* chi, const float2* v, const double* nrm, float2* vw, float2* z) {
  FnoIO io;
  io.chi = chi;
  io.sChi = n;
  io.v = v;
  io.sV = n;
  io.nrm = nrm;
  io.vw = vw;
  io.z = z;
  io.sZ = n;
  return io;
}

// One preconditioned cycle: z_i = P(v_i), w = A z_i, CGS2 step; the caller
// applies x += Z y.
static int precond_cycle(cgm_context* c, const float2* gq, const float* chi, const Precond& P, int m, double* hist,
                         int total) {
  const bool has_pc = P.net || P.tiles;
  KrylovWs& ws = c->ws;
  const long n = ws.n;
  int rc;
  if (c->path != 1 && fused_supported(c->N, m)) {
    // flexible fused lagged-CGS2 steps: z_j = P(u_j / ||u_j||), w = A z_j,
    // one reduction and one basis update per step; u_0 = r (in V_0, norm in
    // st->beta), later norms from B(j-1) in st->unorm
    CGM_HIP(hipMemcpyAsync(ws.U, ws.V, sizeof(float2) * n, hipMemcpyDeviceToDevice, c->stream));
    FusedLaunch L{};
    L.N = c->N; L.m = m; L.ld = ws.m; L.mode = kModeForward; L.chi = chi; L.V = ws.V; L.Z = ws.Z;
    L.T = c->T32; L.tw = c->tw32; L.part = ws.part; L.H = ws.H; L.Rm = ws.R; L.cs = ws.cs;
    L.sn = ws.sn; L.g = ws.g; L.h1p = ws.h1p; L.a1 = ws.a1; L.a2 = ws.a2; L.scal = ws.scal;
    L.st = c->st; L.hist = hist ? hist + total : nullptr;
    L.flex = 1;
    L.partn = ws.partn;
    L.small = fused_small(c, 1);
    for (int j = 0; j <= m; ++j) {
      L.j = j;
      L.Uj = ws.U + (long)(j & 1) * n;
      L.Unext = ws.U + (long)((j + 1) & 1) * n;
      L.Xin = nullptr;
      if (j < m) {
        float2* zj = c->zb + (long)j * n;
        const double* nrm = j == 0 ? &c->st->beta : &c->st->unorm;
        if (has_pc) {
          rc = precond_apply(c, gq, P, c->stream, single_io(n, chi, L.Uj, nrm, nullptr, zj), 1, c->st, 0, c->T32);
        } else {
          hipLaunchKernelGGL(k_scale_copy, dim3(ws.G), dim3(kKB), 0, c->stream, L.Uj, nrm, zj, n, &c->st->done);
          rc = hipGetLastError() == hipSuccess ? CGM_SUCCESS : CGM_ERROR_DEVICE;
        }
        if (!rc)
          rc = launch_pass1<float>(c->plan, c->stream, zj, nullptr, chi, nullptr, kModeForward, c->T32, c->tw32,
                                   &c->st->done);
        if (!rc) rc = launch_pass2<float, float2>(c->plan, c->stream, c->T32, gq, c->tw32, &c->st->done);
        if (rc) return rc;
        L.Xin = zj;
      }
      if ((rc = launch_fused(c->stream, 0, L))) return rc;
      if (j == m) break;
      if ((rc = launch_fused(c->stream, 1, L))) return rc;
    }
    return CGM_SUCCESS;
  }
  for (int i = 0; i < m; ++i) {
    float2* vi = ws.V + (long)i * n;
    float2* zi = c->zb + (long)i * n;
    const double* nrm = (i == 0) ? &c->st->beta : &c->st->hn;
    if (has_pc) {
      rc = precond_apply(c, gq, P, c->stream, single_io(n, chi, vi, nrm, vi, zi), 1, c->st, 0, c->T32);
    } else {
      hipLaunchKernelGGL(k_norm_copy, dim3(ws.G), dim3(kKB), 0, c->stream, vi, nrm, zi, n, &c->st->done);
      rc = hipGetLastError() == hipSuccess ? CGM_SUCCESS : CGM_ERROR_DEVICE;
    }
    if (rc) return rc;
    rc = matvec32(c, gq, chi, zi, nullptr, vi + n, kModeForward, nullptr, nullptr, &c->st->done);
    if (rc) return rc;
    rc = launch_cgs2_step(c->stream, ws, i, c->st, hist ? hist + 1 + total : nullptr);
    if (rc) return rc;
  }
  return CGM_SUCCESS;
}

// The preconditioner must belong to the context and, for the tiles, hold
// maps for the systems (one shared map, or at least `systems`).
static int check_precond(cgm_context* c, const Precond& P, int systems, long chi_stride) {
  CGM_CHECK(!P.net || fno_context(P.net) == c, CGM_ERROR_INVALID_VALUE,
            "the preconditioner was created on another context");
  CGM_CHECK(!P.net || fno_batch(P.net) >= systems, CGM_ERROR_INVALID_VALUE,
            "the preconditioner was created for %d vectors per launch, fewer than batch %d",
            P.net ? fno_batch(P.net) : 0, systems);
  if (P.tiles) {
    CGM_CHECK(tiles_context(P.tiles) == c, CGM_ERROR_INVALID_VALUE, "the tiles were created on another context");
    const int maps = tiles_maps(P.tiles);
    CGM_CHECK(maps >= 1, CGM_ERROR_INVALID_VALUE, "the tile preconditioner has not been factored");
    CGM_CHECK(maps == 1 ? (systems == 1 || chi_stride == 0) : (maps >= systems && chi_stride != 0),
              CGM_ERROR_INVALID_VALUE, "%d factored maps do not match %d systems (chi_stride %ld)", maps, systems,
              chi_stride);
  }
  return CGM_SUCCESS;
}

// ---------------------------------------------------------------------------
// Long-recurrence flexible GMRES (restart > 64): B systems (B = 1 uses the
// context's own state), cycles of up to kMaxRestartLR steps with the whole
// basis in HBM (cgm_longrec.hip).
// ---------------------------------------------------------------------------
struct LrCall {
  cgm_context* c;
  const float2* gq;
  const float* chi;
  long chi_stride;
  const Precond* P;
  bool use_pc;
  DevState* dst;
  long sst;
  int B;
  float2* Tm;
};

static int lr_precond_step(void* vp, int, const float2* u, long su, float2* z, long sz) {
  const LrCall& L = *static_cast<const LrCall*>(vp);
  const long n = (long)L.c->N * L.c->N;
  if (!L.use_pc) return lr_scale(L.c->stream, u, su, z, sz, n, L.dst, L.sst, L.B);
  FnoIO io;
  io.chi = L.chi;
  io.sChi = L.chi_stride;
  io.v = u;
  io.sV = su;
  io.nrm = &L.dst->unorm;
  io.sNrm = L.sst;
  io.z = z;
  io.sZ = sz;
  return precond_apply(L.c, L.gq, *L.P, L.c->stream, io, L.B, L.dst, L.sst, L.Tm);
}

// Strided precision conversions of B vectors of n elements (grid y = vector).
__global__ void k_c64_to_c128(const float2* __restrict__ a, long sa, double2* __restrict__ o, long so, long n) {
  a += (long)blockIdx.y * sa;
  o += (long)blockIdx.y * so;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    o[e] = make_double2(a[e].x, a[e].y);
}
__global__ void k_c128_to_c64(const double2* __restrict__ a, long sa, float2* __restrict__ o, long so, long n) {
  a += (long)blockIdx.y * sa;
  o += (long)blockIdx.y * so;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    o[e] = make_float2((float)a[e].x, (float)a[e].y);
}
// y = b - w (b complex64, w complex128 = A x), rounded to complex64.
__global__ void k_resid_mixed(const float2* __restrict__ b, long sb, const double2* __restrict__ w, long sw,
                              float2* __restrict__ y, long sy, long n) {
  b += (long)blockIdx.y * sb;
  w += (long)blockIdx.y * sw;
  y += (long)blockIdx.y * sy;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    y[e] = make_float2((float)((double)b[e].x - w[e].x), (float)((double)b[e].y - w[e].y));
}
__global__ void k_f32_to_f64(const float* __restrict__ a, long sa, double* __restrict__ o, long so, long n) {
  a += (long)blockIdx.y * sa;
  o += (long)blockIdx.y * so;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) o[e] = a[e];
}

extern "C" int cgm_set_flexible_fp64(cgm_ctx_t c, const cgm_complex128* g64) {
  CGM_CHECK(c, CGM_ERROR_INVALID_VALUE, "null context");
  c->flex_g64 = (const double2*)g64;
  return CGM_SUCCESS;
}

// FP64 buffers of the long-recurrence matvecs for B systems and `maps` maps.
static int ensure_flex64(cgm_context* c, int B, int maps) {
  const long n = (long)c->N * c->N;
  if (c->chi64_len < (long)maps * n) {
    if (c->chi64) CGM_HIP(hipFree(c->chi64));
    c->chi64 = nullptr;
    c->chi64_len = 0;
    CGM_HIP(hipMalloc((void**)&c->chi64, sizeof(double) * (size_t)maps * n));
    c->chi64_len = (long)maps * n;
  }
  if (c->v64_len < (long)B * n) {
    if (c->z64) CGM_HIP(hipFree(c->z64));
    if (c->w64) CGM_HIP(hipFree(c->w64));
    c->z64 = c->w64 = nullptr;
    c->v64_len = 0;
    CGM_HIP(hipMalloc((void**)&c->z64, sizeof(double2) * (size_t)B * n));
    CGM_HIP(hipMalloc((void**)&c->w64, sizeof(double2) * (size_t)B * n));
    c->v64_len = (long)B * n;
  }
  if (c->T64B_n < B) {
    if (c->T64B) CGM_HIP(hipFree(c->T64B));
    c->T64B = nullptr;
    c->T64B_n = 0;
    CGM_HIP(hipMalloc((void**)&c->T64B, sizeof(double2) * (size_t)B * c->N * c->M));
    c->T64B_n = B;
  }
  return CGM_SUCCESS;
}

// y = b - A x (resid) or A x in FP64 for B systems: x (complex64, stride sx)
// is widened into z64, the FP64 passes write w64, which is rounded into y
// (stride sy).  Exact for the complex64 input up to FP64 rounding.
static int matvec64_from32(cgm_context* c, const float2* x, long sx, const float2* b, long sb, float2* y, long sy,
                           bool resid, long chi_stride, int B, const int* done, long sst, bool keep64 = false) {
  const long n = (long)c->N * c->N;
  const dim3 grid((unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024), B);
  hipLaunchKernelGGL(k_c64_to_c128, grid, dim3(256), 0, c->stream, x, sx, c->z64, n, n);
  CGM_HIP(hipGetLastError());
  BatchStrides bs;
  bs.x = n;
  bs.chi = chi_stride ? n : 0;
  bs.y = n;
  bs.T = (long)c->N * c->M;
  bs.st = sst;
  int rc = launch_pass1<double>(c->plan64, c->stream, c->z64, nullptr, c->chi64, nullptr, kModeForward, c->T64B,
                                c->tw64, done, B, bs);
  if (!rc) rc = launch_pass2<double, double2>(c->plan64, c->stream, c->T64B, c->flex_g64, c->tw64, done, B, bs);
  if (!rc)
    rc = launch_pass3<double>(c->plan64, c->stream, c->T64B, c->z64, c->chi64, nullptr, c->w64, kModeForward,
                              c->tw64, done, B, bs);
  if (rc) return rc;
  if (resid)
    hipLaunchKernelGGL(k_resid_mixed, grid, dim3(256), 0, c->stream, b, sb, c->w64, n, y, sy, n);
  else if (keep64)
    return CGM_SUCCESS;  // the caller reads w64
  else
    hipLaunchKernelGGL(k_c128_to_c64, grid, dim3(256), 0, c->stream, c->w64, n, y, sy, n);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

static int lr_matvec_step(void* vp, int, const float2* z, long sz, float2* w, long sw) {
  const LrCall& L = *static_cast<const LrCall*>(vp);
  cgm_context* c = L.c;
  // FP64 products stay in w64, which the step kernels read (c->lr.W64)
  if (c->flex_g64)
    return matvec64_from32(c, z, sz, nullptr, 0, w, sw, false, L.chi_stride, L.B, &L.dst->done, L.sst, true);
  BatchStrides bs;
  bs.x = sz;
  bs.chi = L.chi_stride;
  bs.y = sw;
  bs.T = (long)c->N * c->M;
  bs.st = L.sst;
  const int* done = &L.dst->done;
  int rc = launch_pass1<float>(c->plan, c->stream, z, nullptr, L.chi, nullptr, kModeForward, L.Tm, c->tw32, done, L.B,
                               bs);
  if (!rc) rc = launch_pass2<float, float2>(c->plan, c->stream, L.Tm, L.gq, c->tw32, done, L.B, bs);
  if (!rc)
    rc = launch_pass3<float>(c->plan, c->stream, L.Tm, z, L.chi, nullptr, w, kModeForward, c->tw32, done, L.B, bs);
  return rc;
}

static int ensure_states(cgm_context* c, int B) {
  if (c->stB_n >= B) return CGM_SUCCESS;
  if (c->stB) CGM_HIP(hipFree(c->stB));
  if (c->hstB) CGM_HIP(hipHostFree(c->hstB));
  c->stB = nullptr;
  c->hstB = nullptr;
  c->stB_n = 0;
  CGM_HIP(hipMalloc((void**)&c->stB, sizeof(DevState) * B));
  CGM_HIP(hipHostMalloc((void**)&c->hstB, sizeof(DevState) * B));
  c->stB_n = B;
  return CGM_SUCCESS;
}

// results[k] for k < B; `batched` selects the per-system state array (B may
// be 1 there too).  stall_ratio applies to B = 1 only (the caller checks).
static int lr_solve(cgm_context* c, const float2* gq, const float* chi, long chi_stride, const float2* b, float2* x,
                    const Precond& P, int B, bool batched, const cgm_fgmres_params* p, cgm_fgmres_result* res,
                    double* hist) {
  const long n = (long)c->N * c->N;
  const int m = p->restart;
  int rc = krylov_alloc(c->ws, n, 1, c->stream, B);
  if (!rc) rc = lr_alloc(c->lr, n, m, B);
  if (rc) {
    lr_free(c->lr);
    cgm_set_last_error("long-recurrence workspace of %zu MiB for %d systems, restart %d: allocation failed",
                       lr_bytes(n, m, B) >> 20, B, m);
    return CGM_ERROR_ALLOC;
  }
  DevState* dst = c->st;
  DevState* hst = c->hst;
  long sst = 0;
  float2* Tm = c->T32;
  if (batched) {
    if ((rc = ensure_states(c, B))) return rc;
    if ((rc = ensure_T32B(c, B))) return rc;
    dst = c->stB;
    hst = c->hstB;
    sst = (long)sizeof(DevState);
    Tm = c->T32B;
  }
  if (P.net && P.tiles && (rc = ensure_pc(c, B))) return rc;
  if (c->flex_g64) {
    const int maps = chi_stride ? B : 1;
    if ((rc = ensure_flex64(c, B, maps))) return rc;
    const dim3 grid((unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024), maps);
    hipLaunchKernelGGL(k_f32_to_f64, grid, dim3(256), 0, c->stream, chi, chi_stride, c->chi64, n, n);
    CGM_HIP(hipGetLastError());
  }
  c->lr.W64 = c->flex_g64 ? c->w64 : nullptr;
  const long hstride = (long)m * (p->max_cycles > 0 ? p->max_cycles : 1) + 1;
  CGM_HIP(hipMemsetAsync(dst, 0, sizeof(DevState) * (batched ? B : 1), c->stream));
  if (hist) CGM_HIP(hipMemsetAsync(hist, 0, sizeof(double) * hstride * B, c->stream));
  if ((rc = launch_norm_init(c->stream, c->ws, b, dst, 1, p->tol, n, sst, B))) return rc;
  if (p->x0_zero) CGM_HIP(hipMemsetAsync(x, 0, sizeof(float2) * n * B, c->stream));
  LrCall L{c, gq, chi, chi_stride, &P, P.net || P.tiles, dst, sst, B, Tm};
  const LrStep pre{lr_precond_step, &L};
  const LrStep mv{lr_matvec_step, &L};
  BatchStrides rs;  // r = b - A x into u_0 of every system
  rs.x = n;
  rs.chi = chi_stride;
  rs.b = n;
  rs.y = 2 * n;
  rs.T = (long)c->N * c->M;
  rs.st = sst;
  std::vector<int> fin(B, 0);
  std::vector<double> beta_prev(B, 0.0);
  int cycles = 0;
  bool last_pc = false;
  for (;;) {
    if (p->x0_zero && cycles == 0) {
      rc = copy_strided(c->stream, b, n, c->lr.U, 2 * n, n, B);
    } else if (c->flex_g64) {
      rc = matvec64_from32(c, x, n, b, n, c->lr.U, 2 * n, true, chi_stride, B, nullptr, sst);
    } else {
      rc = launch_pass1<float>(c->plan, c->stream, x, nullptr, chi, nullptr, kModeResidual, Tm, c->tw32, nullptr, B, rs);
      if (!rc) rc = launch_pass2<float, float2>(c->plan, c->stream, Tm, gq, c->tw32, nullptr, B, rs);
      if (!rc)
        rc = launch_pass3<float>(c->plan, c->stream, Tm, x, chi, b, c->lr.U, kModeResidual, c->tw32, nullptr, B, rs);
    }
    if (!rc) rc = launch_norm_init(c->stream, c->ws, c->lr.U, dst, 0, 0.0, 2 * n, sst, B);
    if (rc) return rc;
    CGM_HIP(hipMemcpyAsync(hst, dst, sizeof(DevState) * (batched ? B : 1), hipMemcpyDeviceToHost, c->stream));
    CGM_HIP(hipStreamSynchronize(c->stream));
    int active = 0;
    for (int k = 0; k < B; ++k) {
      const DevState& h = *boff(hst, (long)k * sst);
      cgm_fgmres_result& r = res[k];
      if (cycles == 0) {
        r.beta0 = h.beta;
        r.bnorm = h.bnorm;
        if (hist)
          CGM_HIP(hipMemcpyAsync(hist + k * hstride, &boff(hst, (long)k * sst)->beta, sizeof(double),
                                 hipMemcpyHostToDevice, c->stream));
      }
      if (fin[k]) continue;
      r.rel_residual = h.bnorm > 0 ? h.beta / h.bnorm : 0.0;
      if (h.status) {
        fin[k] = 1;
        continue;
      }
      if (h.bnorm == 0.0 || h.beta <= p->tol * h.bnorm) {
        r.converged = 1;
        fin[k] = 1;
        continue;
      }
      ++active;
    }
    if (!active || cycles >= p->max_cycles) break;
    // the stall rule of cgm_fgmres (one system): a preconditioned cycle that
    // left more than stall_ratio of the residual hands over to the identity
    if (B == 1 && L.use_pc && last_pc && p->stall_ratio > 0.0 && hst->beta > p->stall_ratio * beta_prev[0]) {
      L.use_pc = false;
      res[0].fell_back = 1;
    }
    beta_prev[0] = hst->beta;
    last_pc = L.use_pc;
    if ((rc = lr_cycle(c->lr, c->stream, dst, hst, sst, B, m, hist ? hist + (long)cycles * m : nullptr, hstride, x, n,
                       pre, mv, c->lrt.on ? &c->lrt : nullptr)))
      return rc;
    CGM_HIP(hipMemcpyAsync(hst, dst, sizeof(DevState) * (batched ? B : 1), hipMemcpyDeviceToHost, c->stream));
    CGM_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < B; ++k) {
      if (fin[k]) continue;
      const DevState& h = *boff(hst, (long)k * sst);
      res[k].iterations += h.k_last;
      res[k].cycles += 1;
      if (L.use_pc) res[k].precond_cycles += 1;
    }
    ++cycles;
  }
  int any_nonfinite = 0, all_conv = 1;
  for (int k = 0; k < B; ++k) {
    if (boff(hst, (long)k * sst)->status) any_nonfinite = 1;
    if (!res[k].converged) all_conv = 0;
  }
  if (any_nonfinite) {
    cgm_set_last_error("flexible GMRES: non-finite value in at least one system");
    return CGM_ERROR_NONFINITE;
  }
  return all_conv ? CGM_SUCCESS : CGM_ERROR_NOT_CONVERGED;
}

static int fgmres_core(cgm_context* c, const cgm_complex64* gamma, const float* chi, const cgm_complex64* b,
                       cgm_complex64* x, const Precond& P, const cgm_fgmres_params* p, cgm_fgmres_result* res,
                       double* hist);

extern "C" int cgm_fgmres(cgm_ctx_t c, const cgm_complex64* gamma, const float* chi, const cgm_complex64* b,
                          cgm_complex64* x, cgm_fno_t net, const cgm_fgmres_params* p, cgm_fgmres_result* res,
                          double* hist) {
  Precond P;
  P.net = net;
  return fgmres_core(c, gamma, chi, b, x, P, p, res, hist);
}

extern "C" int cgm_fgmres_pc(cgm_ctx_t c, const cgm_complex64* gamma, const float* chi, const cgm_complex64* b,
                             cgm_complex64* x, const cgm_precond* pc, const cgm_fgmres_params* p,
                             cgm_fgmres_result* res, double* hist) {
  Precond P;
  if (pc) {
    P.net = pc->net;
    P.tiles = pc->tiles;
  }
  return fgmres_core(c, gamma, chi, b, x, P, p, res, hist);
}

static int fgmres_core(cgm_context* c, const cgm_complex64* gamma, const float* chi, const cgm_complex64* b,
                       cgm_complex64* x, const Precond& P, const cgm_fgmres_params* p, cgm_fgmres_result* res,
                       double* hist) {
  CGM_CHECK(c && gamma && chi && b && x && res && p, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(p->restart >= 1 && p->restart <= kMaxRestartLR, CGM_ERROR_INVALID_VALUE, "restart %d outside [1, %d]",
            p->restart, kMaxRestartLR);
  CGM_CHECK(p->max_cycles >= 0, CGM_ERROR_INVALID_VALUE, "max_cycles < 0");
  CGM_CHECK(p->tol >= 0 && isfinite(p->tol), CGM_ERROR_INVALID_VALUE, "bad tolerance");
  int rc = check_precond(c, P, 1, 0);
  if (rc) return rc;
  CGM_HIP(hipSetDevice(c->device));
  memset(res, 0, sizeof(*res));
  if (p->restart > kFusedMaxRestart)
    return lr_solve(c, (const float2*)gamma, chi, 0, (const float2*)b, (float2*)x, P, 1, false, p, res, hist);
  const long n = (long)c->N * c->N;
  const int m = p->restart;
  rc = krylov_alloc(c->ws, n, m, c->stream);
  if (rc) return rc;
  if ((rc = ensure_zb(c, n * m))) return rc;
  if (P.net && P.tiles && (rc = ensure_pc(c, 1))) return rc;
  KrylovWs& ws = c->ws;
  const float2* bb = (const float2*)b;
  float2* xx = (float2*)x;
  CGM_HIP(hipMemsetAsync(&c->st->status, 0, sizeof(int), c->stream));
  if ((rc = launch_norm_init(c->stream, ws, bb, c->st, 1, p->tol))) return rc;
  if (p->x0_zero) CGM_HIP(hipMemsetAsync(xx, 0, sizeof(float2) * n, c->stream));
  int total = 0, cycles = 0;
  bool precond = true;
  bool last_precond = false;
  double beta_prev = 0.0;
  for (;;) {
    if (p->x0_zero && cycles == 0) {
      CGM_HIP(hipMemcpyAsync(ws.V, bb, sizeof(float2) * n, hipMemcpyDeviceToDevice, c->stream));
    } else {
      rc = matvec32(c, (const float2*)gamma, chi, xx, bb, ws.V, kModeResidual, nullptr, nullptr, nullptr);
      if (rc) return rc;
    }
    if ((rc = launch_norm_init(c->stream, ws, ws.V, c->st, 0, 0.0))) return rc;
    if ((rc = read_state(c))) return rc;
    const DevState& h = *c->hst;
    if (cycles == 0) {
      res->beta0 = h.beta;
      res->bnorm = h.bnorm;
      if (hist) CGM_HIP(hipMemcpyAsync(hist, &h.beta, sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    res->rel_residual = h.bnorm > 0 ? h.beta / h.bnorm : 0.0;
    if (h.status) {
      cgm_set_last_error("cgm_fgmres: non-finite residual");
      res->iterations = total;
      res->cycles = cycles;
      return CGM_ERROR_NONFINITE;
    }
    if (h.bnorm == 0.0 || h.beta <= p->tol * h.bnorm) { res->converged = 1; break; }
    if (cycles >= p->max_cycles) break;
    // the true residual decides on the preconditioner: a cycle that did not cut
    // it below stall_ratio hands the rest of the solve to plain GMRES
    if (precond && last_precond && p->stall_ratio > 0.0 && h.beta > p->stall_ratio * beta_prev) {
      precond = false;
      res->fell_back = 1;
    }
    beta_prev = h.beta;
    if (precond) {
      if ((rc = precond_cycle(c, (const float2*)gamma, chi, P, m, hist, total))) return rc;
      if ((rc = launch_backsub_update(c->stream, ws, c->st, xx, c->zb))) return rc;
      res->precond_cycles += 1;
    } else {
      if ((rc = plain_cycle(c, (const float2*)gamma, chi, kModeForward, m, hist, total))) return rc;
      if ((rc = launch_backsub_update(c->stream, ws, c->st, xx))) return rc;
    }
    last_precond = precond;
    if ((rc = read_state(c))) return rc;
    total += c->hst->k;
    ++cycles;
    if (c->hst->status) {
      cgm_set_last_error("cgm_fgmres: non-finite value in the Arnoldi process");
      res->iterations = total;
      res->cycles = cycles;
      return CGM_ERROR_NONFINITE;
    }
  }
  res->iterations = total;
  res->cycles = cycles;
  return res->converged ? CGM_SUCCESS : CGM_ERROR_NOT_CONVERGED;
}

// ---------------------------------------------------------------------------
// Batched flexible GMRES: B systems of one grid size in lockstep, one network
// launch per step for all of them
// ---------------------------------------------------------------------------
static int fgmres_batched_core(cgm_context* c, const cgm_complex64* gamma, const float* chi, long chi_stride,
                               const cgm_complex64* b, cgm_complex64* x, const Precond& P, int batch,
                               const cgm_fgmres_params* p, cgm_fgmres_result* results, double* hist);

extern "C" int cgm_fgmres_batched(cgm_ctx_t c, const cgm_complex64* gamma, const float* chi, long chi_stride,
                                  const cgm_complex64* b, cgm_complex64* x, cgm_fno_t net, int batch,
                                  const cgm_fgmres_params* p, cgm_fgmres_result* results, double* hist) {
  Precond P;
  P.net = net;
  return fgmres_batched_core(c, gamma, chi, chi_stride, b, x, P, batch, p, results, hist);
}

extern "C" int cgm_fgmres_pc_batched(cgm_ctx_t c, const cgm_complex64* gamma, const float* chi, long chi_stride,
                                     const cgm_complex64* b, cgm_complex64* x, const cgm_precond* pc, int batch,
                                     const cgm_fgmres_params* p, cgm_fgmres_result* results, double* hist) {
  Precond P;
  if (pc) {
    P.net = pc->net;
    P.tiles = pc->tiles;
  }
  return fgmres_batched_core(c, gamma, chi, chi_stride, b, x, P, batch, p, results, hist);
}

static int fgmres_batched_core(cgm_context* c, const cgm_complex64* gamma, const float* chi, long chi_stride,
                               const cgm_complex64* b, cgm_complex64* x, const Precond& P, int batch,
                               const cgm_fgmres_params* p, cgm_fgmres_result* results, double* hist) {
  CGM_CHECK(c && gamma && chi && b && x && results && p, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(batch >= 1, CGM_ERROR_INVALID_VALUE, "batch %d < 1", batch);
  CGM_CHECK(chi_stride == 0 || chi_stride >= (long)c->N * c->N, CGM_ERROR_INVALID_VALUE,
            "chi_stride must be 0 (shared map) or >= N*N");
  CGM_CHECK(p->restart >= 1 && p->restart <= kMaxRestartLR, CGM_ERROR_INVALID_VALUE, "restart %d outside [1, %d]",
            p->restart, kMaxRestartLR);
  CGM_CHECK(p->max_cycles >= 0, CGM_ERROR_INVALID_VALUE, "max_cycles < 0");
  CGM_CHECK(p->tol >= 0 && isfinite(p->tol), CGM_ERROR_INVALID_VALUE, "bad tolerance");
  CGM_CHECK(p->stall_ratio == 0.0, CGM_ERROR_UNSUPPORTED,
            "cgm_fgmres_batched has no plain-GMRES fallback: stall_ratio must be 0");
  CGM_CHECK(p->restart > kFusedMaxRestart || fused_supported(c->N, p->restart), CGM_ERROR_UNSUPPORTED,
            "cgm_fgmres_batched with restart <= 64 needs a power-of-two N in [64, 2048]");
  {
    const int rcp = check_precond(c, P, batch, chi_stride);
    if (rcp) return rcp;
  }
  const bool has_pc = P.net || P.tiles;
  CGM_HIP(hipSetDevice(c->device));
  memset(results, 0, sizeof(cgm_fgmres_result) * batch);
  if (p->restart > kFusedMaxRestart)
    return lr_solve(c, (const float2*)gamma, chi, chi_stride, (const float2*)b, (float2*)x, P, batch, true, p,
                    results, hist);
  const long n = (long)c->N * c->N;
  const int m = p->restart;
  const int B = batch;
  const long hstride = (long)m * (p->max_cycles > 0 ? p->max_cycles : 1) + 1;
  int rc = krylov_alloc(c->ws, n, m, c->stream, B);
  if (rc) return rc;
  KrylovWs& ws = c->ws;
  if (c->stB_n < B) {
    if (c->stB) CGM_HIP(hipFree(c->stB));
    if (c->hstB) CGM_HIP(hipHostFree(c->hstB));
    c->stB = nullptr;
    c->hstB = nullptr;
    c->stB_n = 0;
    CGM_HIP(hipMalloc((void**)&c->stB, sizeof(DevState) * B));
    CGM_HIP(hipHostMalloc((void**)&c->hstB, sizeof(DevState) * B));
    c->stB_n = B;
  }
  if ((rc = ensure_T32B(c, B))) return rc;
  if ((rc = ensure_zb(c, ws.sBasis * B))) return rc;  // Z of system k at zb + k * sBasis
  if (P.net && P.tiles && (rc = ensure_pc(c, B))) return rc;
  const long NM = (long)c->N * c->M;
  const long sst = (long)sizeof(DevState);
  const float2* gq = (const float2*)gamma;
  const float2* bb = (const float2*)b;
  float2* xx = (float2*)x;
  DevState* stB = c->stB;
  CGM_HIP(hipMemsetAsync(stB, 0, sizeof(DevState) * B, c->stream));
  if (hist) CGM_HIP(hipMemsetAsync(hist, 0, sizeof(double) * hstride * B, c->stream));
  if ((rc = launch_norm_init(c->stream, ws, bb, stB, 1, p->tol, n, sst, B))) return rc;
  if (p->x0_zero) CGM_HIP(hipMemsetAsync(xx, 0, sizeof(float2) * n * B, c->stream));
  BatchStrides rs;  // residual r = b - A x of every system into its V_0
  rs.x = n;
  rs.chi = chi_stride;
  rs.b = n;
  rs.y = ws.sBasis;
  rs.T = NM;
  rs.st = sst;
  BatchStrides zs = rs;  // the matvec of z_j
  zs.x = ws.sBasis;
  auto residual = [&](hipStream_t s) -> int {
    int r = launch_pass1<float>(c->plan, s, xx, nullptr, chi, nullptr, kModeResidual, c->T32B, c->tw32, nullptr, B, rs);
    if (!r) r = launch_pass2<float, float2>(c->plan, s, c->T32B, gq, c->tw32, nullptr, B, rs);
    if (!r) r = launch_pass3<float>(c->plan, s, c->T32B, xx, chi, bb, ws.V, kModeResidual, c->tw32, nullptr, B, rs);
    if (!r) r = launch_norm_init(s, ws, ws.V, stB, 0, 0.0, ws.sBasis, sst, B);
    return r;
  };
  // one flexible restart cycle of every live system, x += Z y, the next
  // residual and its norm
  auto enqueue_cycle = [&](hipStream_t s, int cyc) -> int {
    int r = copy_strided(s, ws.V, ws.sBasis, ws.U, 2 * n, n, B);
    if (r) return r;
    FusedLaunch L{};
    L.N = c->N; L.m = m; L.ld = ws.m; L.mode = kModeForward; L.chi = chi; L.V = ws.V; L.Z = ws.Z;
    L.T = c->T32B; L.tw = c->tw32; L.part = ws.part; L.H = ws.H; L.Rm = ws.R; L.cs = ws.cs;
    L.sn = ws.sn; L.g = ws.g; L.h1p = ws.h1p; L.a1 = ws.a1; L.a2 = ws.a2; L.scal = ws.scal;
    L.st = stB; L.hist = hist ? hist + (long)cyc * m : nullptr;
    L.B = B; L.sChi = chi_stride; L.sBasis = ws.sBasis; L.sU = 2 * n; L.sVec = n; L.sT = NM;
    L.sPart = ws.sPart; L.sSmall = ws.sSmall; L.sSt = sst; L.sHist = hstride;
    L.flex = 1;
    L.partn = ws.partn;
    L.sPartn = ws.sPartn;
    L.sXin = ws.sBasis;
    L.small = fused_small(c, B);
    for (int j = 0; j <= m; ++j) {
      L.j = j;
      L.Uj = ws.U + (long)(j & 1) * n;
      L.Unext = ws.U + (long)((j + 1) & 1) * n;
      L.Xin = nullptr;
      if (j < m) {
        float2* zj = c->zb + (long)j * n;
        const double* nrm = j == 0 ? &stB->beta : &stB->unorm;
        if (has_pc) {
          FnoIO io;
          io.chi = chi;
          io.sChi = chi_stride;
          io.v = L.Uj;
          io.sV = 2 * n;
          io.nrm = nrm;
          io.sNrm = sst;
          io.z = zj;
          io.sZ = ws.sBasis;
          r = precond_apply(c, gq, P, s, io, B, stB, sst, c->T32B);
        } else {
          r = copy_strided(s, L.Uj, 2 * n, zj, ws.sBasis, n, B);  // z = u (scaled below)
          if (!r) {
            hipLaunchKernelGGL(k_scale_inplace, dim3(ws.G, B), dim3(kKB), 0, s, zj, ws.sBasis, nrm, sst, n);
            r = hipGetLastError() == hipSuccess ? CGM_SUCCESS : CGM_ERROR_DEVICE;
          }
        }
        if (!r)
          r = launch_pass1<float>(c->plan, s, zj, nullptr, chi, nullptr, kModeForward, c->T32B, c->tw32, &stB->done,
                                  B, zs);
        if (!r) r = launch_pass2<float, float2>(c->plan, s, c->T32B, gq, c->tw32, &stB->done, B, zs);
        if (r) return r;
        L.Xin = zj;
      }
      if ((r = launch_fused(s, 0, L))) return r;
      if (j == m) break;
      if ((r = launch_fused(s, 1, L))) return r;
    }
    if ((r = launch_backsub_update(s, ws, stB, xx, c->zb, n, sst, B))) return r;
    return residual(s);
  };
  // first residual, eagerly on the context stream
  if (p->x0_zero) {
    if ((rc = copy_strided(c->stream, bb, n, ws.V, ws.sBasis, n, B))) return rc;
    if ((rc = launch_norm_init(c->stream, ws, ws.V, stB, 0, 0.0, ws.sBasis, sst, B))) return rc;
  } else if ((rc = residual(c->stream))) {
    return rc;
  }
  if ((rc = read_states(c, B, c->stream))) return rc;
  const bool graph = c->use_graphs && !hist;
  hipStream_t cs = c->stream;
  hipGraphExec_t exec = nullptr;
  if (graph) {
    if (!c->gstream) CGM_HIP(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
    cs = c->gstream;
  }
  std::vector<int> done(B, 0);
  int cycles = 0;
  for (;;) {
    int active = 0;
    for (int k = 0; k < B; ++k) {
      const DevState& h = c->hstB[k];
      cgm_fgmres_result& r = results[k];
      if (cycles == 0) {
        r.beta0 = h.beta;
        r.bnorm = h.bnorm;
        if (hist) CGM_HIP(hipMemcpyAsync(hist + k * hstride, &c->hstB[k].beta, sizeof(double),
                                         hipMemcpyHostToDevice, c->stream));
      } else if (!done[k]) {
        r.iterations += h.k_last;
        r.cycles += 1;
        r.precond_cycles += has_pc ? 1 : 0;
      }
      if (done[k]) continue;
      r.rel_residual = h.bnorm > 0 ? h.beta / h.bnorm : 0.0;
      if (h.status) {
        done[k] = 1;
        continue;
      }
      if (h.bnorm == 0.0 || h.beta <= p->tol * h.bnorm) {
        r.converged = 1;
        done[k] = 1;
        continue;
      }
      ++active;
    }
    if (!active || cycles >= p->max_cycles) break;
    if (graph) {
      if (!exec) {
        hipGraph_t gr = nullptr;
        CGM_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        rc = enqueue_cycle(cs, 0);
        const hipError_t ec = hipStreamEndCapture(cs, &gr);
        if (rc) {
          if (gr) (void)hipGraphDestroy(gr);
          return rc;
        }
        CGM_HIP(ec);
        const hipError_t ei = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        CGM_HIP(ei);
      }
      const hipError_t el = hipGraphLaunch(exec, cs);
      if (el != hipSuccess) {
        (void)hipGraphExecDestroy(exec);
        CGM_HIP(el);
      }
    } else if ((rc = enqueue_cycle(cs, cycles))) {
      return rc;
    }
    rc = read_states(c, B, cs);
    if (rc) {
      if (exec) (void)hipGraphExecDestroy(exec);
      return rc;
    }
    ++cycles;
  }
  if (exec) CGM_HIP(hipGraphExecDestroy(exec));
  int any_nonfinite = 0, all_conv = 1;
  for (int k = 0; k < B; ++k) {
    if (c->hstB[k].status) any_nonfinite = 1;
    if (!results[k].converged) all_conv = 0;
  }
  if (any_nonfinite) {
    cgm_set_last_error("cgm_fgmres_batched: non-finite value in at least one system");
    return CGM_ERROR_NONFINITE;
  }
  return all_conv ? CGM_SUCCESS : CGM_ERROR_NOT_CONVERGED;
}

// ---------------------------------------------------------------------------
// FP64 refinement
// ---------------------------------------------------------------------------
__global__ void k_to_float_scaled(const double2* __restrict__ r, float2* __restrict__ o, long n,
                                  const DevState* st) {
  const double s = st->norm_out > 0 ? 1.0 / st->norm_out : 0.0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    o[e] = make_float2((float)(r[e].x * s), (float)(r[e].y * s));
}

__global__ void k_axpy_mixed(double2* __restrict__ x, const float2* __restrict__ d, long n,
                             const DevState* st) {
  const double s = st->norm_out;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    x[e].x += s * (double)d[e].x;
    x[e].y += s * (double)d[e].y;
  }
}

__global__ void k_d2f(const double* __restrict__ a, float* __restrict__ o, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    o[e] = (float)a[e];
}

// FP64 refinement; the correction is solved by complex64 GMRES(inner_restart)
// (P null) or by flexible GMRES with the preconditioner P (restart up to
// kMaxRestartLR: the long-recurrence path above 64).
static int refine_core(cgm_context* c, const cgm_complex64* g32, const cgm_complex128* g64, const double* chi,
                       const cgm_complex128* b, cgm_complex128* x, const Precond* P, const cgm_refine_params* p,
                       cgm_refine_result* res) {
  CGM_CHECK(c && g32 && g64 && chi && b && x && p && res, CGM_ERROR_INVALID_VALUE, "null argument");
  const int maxr = P ? kMaxRestartLR : kMaxRestart;
  CGM_CHECK(p->inner_restart >= 1 && p->inner_restart <= maxr && p->max_outer >= 0, CGM_ERROR_INVALID_VALUE,
            "bad refinement parameters (inner restart %d outside [1, %d])", p->inner_restart, maxr);
  if (P) {
    const int rcp = check_precond(c, *P, 1, 0);
    if (rcp) return rcp;
  }
  CGM_HIP(hipSetDevice(c->device));
  memset(res, 0, sizeof(*res));
  const long n = (long)c->N * c->N;
  if (!c->v64a) CGM_HIP(hipMalloc((void**)&c->v64a, sizeof(double2) * n));
  if (!c->chi32) CGM_HIP(hipMalloc((void**)&c->chi32, sizeof(float) * n));
  if (!c->vec_tmp) CGM_HIP(hipMalloc((void**)&c->vec_tmp, sizeof(float2) * n));
  float2* d32 = nullptr;
  CGM_HIP(hipMalloc((void**)&d32, sizeof(float2) * n));
  int rc = krylov_alloc(c->ws, n, P ? 1 : p->inner_restart, c->stream);
  if (rc) { (void)hipFree(d32); return rc; }
  const int blocks = 1024;
  hipLaunchKernelGGL(k_d2f, dim3(blocks), dim3(256), 0, c->stream, chi, c->chi32, n);
  // ||b|| in FP64
  rc = launch_norm64(c->stream, (const double2*)b, n, c->ws.partn, c->ws.G, c->st);
  if (!rc) rc = read_state(c);
  const double bn = c->hst->norm_out;
  cgm_gmres_params ip{p->inner_restart, p->inner_max_cycles, p->inner_tol, 0, 1};
  cgm_fgmres_params fp{p->inner_restart, p->inner_max_cycles, p->inner_tol, 1, 0.0};
  for (int it = 0; !rc && it <= p->max_outer; ++it) {
    rc = matvec64(c, (const double2*)g64, chi, (const double2*)x, (const double2*)b, c->v64a,
                  kModeResidual);
    if (rc) break;
    if ((rc = launch_norm64(c->stream, c->v64a, n, c->ws.partn, c->ws.G, c->st))) break;
    if ((rc = read_state(c))) break;
    const double rn = c->hst->norm_out;
    res->rel_residual = bn > 0 ? rn / bn : 0.0;
    if (!isfinite(rn)) { rc = CGM_ERROR_NONFINITE; cgm_set_last_error("refine: non-finite residual"); break; }
    if (rn <= p->outer_tol * bn) { res->converged = 1; break; }
    if (it == p->max_outer) break;
    hipLaunchKernelGGL(k_to_float_scaled, dim3(blocks), dim3(256), 0, c->stream, c->v64a,
                       c->vec_tmp, n, c->st);
    int irc, iters;
    if (P) {
      cgm_fgmres_result ir;
      memset(&ir, 0, sizeof(ir));
      irc = fgmres_core(c, g32, c->chi32, (const cgm_complex64*)c->vec_tmp, (cgm_complex64*)d32, *P, &fp, &ir,
                        nullptr);
      iters = ir.iterations;
    } else {
      cgm_gmres_result ir;
      memset(&ir, 0, sizeof(ir));
      irc = gmres_core(c, (const float2*)g32, c->chi32, c->vec_tmp, d32, &ip, &ir, nullptr);
      iters = ir.iterations;
    }
    if (irc && irc != CGM_ERROR_NOT_CONVERGED) { rc = irc; break; }
    res->inner_iterations += iters;
    res->outer += 1;
    // the inner solve overwrote norm_out; recompute ||r|| for the scale
    if ((rc = launch_norm64(c->stream, c->v64a, n, c->ws.partn, c->ws.G, c->st))) break;
    hipLaunchKernelGGL(k_axpy_mixed, dim3(blocks), dim3(256), 0, c->stream, (double2*)x, d32, n,
                       c->st);
  }
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(d32);
  if (rc) return rc;
  return res->converged ? CGM_SUCCESS : CGM_ERROR_NOT_CONVERGED;
}

extern "C" int cgm_gmres_refine(cgm_ctx_t c, const cgm_complex64* g32, const cgm_complex128* g64,
                                const double* chi, const cgm_complex128* b, cgm_complex128* x,
                                const cgm_refine_params* p, cgm_refine_result* res) {
  return refine_core(c, g32, g64, chi, b, x, nullptr, p, res);
}

extern "C" int cgm_gmres_refine_pc(cgm_ctx_t c, const cgm_complex64* g32, const cgm_complex128* g64,
                                   const double* chi, const cgm_complex128* b, cgm_complex128* x,
                                   const cgm_precond* pc, const cgm_refine_params* p, cgm_refine_result* res) {
  CGM_CHECK(pc, CGM_ERROR_INVALID_VALUE, "null preconditioner");
  Precond P;
  P.net = pc->net;
  P.tiles = pc->tiles;
  return refine_core(c, g32, g64, chi, b, x, &P, p, res);
}

// ---------------------------------------------------------------------------
// diagnostics: per-kernel durations of one Arnoldi step
// ---------------------------------------------------------------------------
extern "C" int cgm_profile_step(cgm_ctx_t c, const cgm_complex64* gamma, const float* chi,
                                const cgm_complex64* x, int restart, int reps, float* ms) {
  CGM_CHECK(c && gamma && chi && x && ms && reps > 0, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(restart >= 1 && restart <= kMaxRestart, CGM_ERROR_INVALID_VALUE, "bad restart");
  CGM_HIP(hipSetDevice(c->device));
  const long n = (long)c->N * c->N;
  int rc = krylov_alloc(c->ws, n, restart, c->stream);
  if (rc) return rc;
  KrylovWs& ws = c->ws;
  const int i = restart / 2;  // mid-cycle step: basis of restart/2 + 1 vectors
  // orthonormal basis e_{j*stride} and w = e_{(i+1)*stride}: every projection
  // is zero, so repeated launches leave the data unchanged (the traffic does
  // not depend on the values)
  CGM_HIP(hipMemsetAsync(ws.V, 0, sizeof(float2) * n * (i + 2), c->stream));
  const float2 one = make_float2(1.f, 0.f);
  const long stride = n / (i + 2);
  for (int j = 0; j <= i + 1; ++j)
    CGM_HIP(hipMemcpy(ws.V + (long)j * n + j * stride, &one, sizeof(float2), hipMemcpyHostToDevice));
  // a state whose tolerance can never be met keeps every launch live
  DevState init{};
  init.tol_abs = -1.0;
  init.hn = 1.0;
  init.beta = 1.0;
  CGM_HIP(hipMemcpyAsync(c->st, &init, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
  std::vector<double2> g1(ws.m + 1, make_double2(1.0, 0.0));
  CGM_HIP(hipMemcpyAsync(ws.g, g1.data(), sizeof(double2) * (ws.m + 1), hipMemcpyHostToDevice, c->stream));
  hipEvent_t e0, e1;
  CGM_HIP(hipEventCreate(&e0));
  CGM_HIP(hipEventCreate(&e1));
  if (!c->vec_tmp) CGM_HIP(hipMalloc((void**)&c->vec_tmp, sizeof(float2) * n));
  for (int k = 0; k < 6 && !rc; ++k) {
    for (int r = -2; r < reps && !rc; ++r) {  // two untimed warm-up launches
      if (r == 0) CGM_HIP(hipEventRecord(e0, c->stream));
      switch (k) {
        case 0: rc = launch_pass1<float>(c->plan, c->stream, (const float2*)x, nullptr, chi, nullptr, kModeForward, c->T32, c->tw32, nullptr); break;
        case 1: rc = launch_pass2<float, float2>(c->plan, c->stream, c->T32, (const float2*)gamma, c->tw32, nullptr); break;
        case 2: rc = launch_pass3<float>(c->plan, c->stream, c->T32, (const float2*)x, chi, nullptr, c->vec_tmp, kModeForward, c->tw32, nullptr); break;
        default: rc = launch_cgs2_step(c->stream, ws, i, c->st, nullptr, k - 3); break;
      }
    }
    CGM_HIP(hipEventRecord(e1, c->stream));
    CGM_HIP(hipEventSynchronize(e1));
    CGM_HIP(hipMemcpyAsync(c->st, &init, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
    float t = 0.f;
    CGM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[k] = t / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// ---------------------------------------------------------------------------
// memory helpers
// ---------------------------------------------------------------------------
extern "C" int cgm_malloc(void** p, size_t bytes) {
  CGM_CHECK(p, CGM_ERROR_INVALID_VALUE, "null pointer");
  if (hipMalloc(p, bytes) != hipSuccess) {
    cgm_set_last_error("hipMalloc of %zu bytes failed", bytes);
    return CGM_ERROR_ALLOC;
  }
  return CGM_SUCCESS;
}
extern "C" int cgm_free(void* p) {
  CGM_HIP(hipFree(p));
  return CGM_SUCCESS;
}
extern "C" int cgm_memcpy_h2d(void* d, const void* s, size_t bytes) {
  CGM_HIP(hipMemcpy(d, s, bytes, hipMemcpyHostToDevice));
  return CGM_SUCCESS;
}
extern "C" int cgm_memcpy_d2h(void* d, const void* s, size_t bytes) {
  CGM_HIP(hipMemcpy(d, s, bytes, hipMemcpyDeviceToHost));
  return CGM_SUCCESS;
}
