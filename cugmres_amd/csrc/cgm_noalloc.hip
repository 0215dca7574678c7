// cgm_matvec_batched_noalloc: the batched FFT matvec for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// cgm_matvec_batched keeps the N x M FFT intermediates of its vectors in a
// context buffer that it grows with hipMalloc when a larger batch arrives.
// This variant runs the same three pass kernels on a fixed array of
// kScratchBytes in device memory instead, so it allocates nothing: the
// array is a file-scope __device__ variable (one copy per GPU, alive until
// the process ends), and its device address is taken once per device with
// hipGetSymbolAddress when a context is created on that device.  Every call
// on a GPU shares that GPU's array; calls that overlap on one GPU (several
// streams or host threads) race on it.  Nothing in it needs clearing: pass 1
// writes every element pass 2 reads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cgm_internal.h"

using namespace cgm;

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB: 16 vectors at N = 512, 4 at N = 1024
constexpr int kMaxDevices = 64;

__device__ __align__(256) float2 scratch[kScratchBytes / sizeof(float2)];

void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

namespace cgm {

// Device address of this device's scratch array (the caller has made
// `device` current).  Called by cgm_create.
int noalloc_init(int device) {
  CGM_CHECK(device >= 0 && device < kMaxDevices, CGM_ERROR_UNSUPPORTED, "device %d beyond %d", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return CGM_SUCCESS;
  void* p = nullptr;
  CGM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  CGM_CHECK(p != nullptr, CGM_ERROR_DEVICE, "hipGetSymbolAddress returned a null scratch address");
  g_scratch_ptr[device] = p;
  return CGM_SUCCESS;
}

}  // namespace cgm

extern "C" size_t cgm_noalloc_scratch_bytes(void) { return kScratchBytes; }

extern "C" int cgm_matvec_batched_noalloc(cgm_ctx_t c, const cgm_complex64* gamma, const float* chi,
                                          long chi_stride, const cgm_complex64* x, cgm_complex64* y, int batch,
                                          int mode) {
  CGM_CHECK(c && gamma && x && y, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(batch >= 1, CGM_ERROR_INVALID_VALUE, "batch %d < 1", batch);
  CGM_CHECK(mode >= 0 && mode <= 2, CGM_ERROR_INVALID_VALUE, "mode %d", mode);
  CGM_CHECK(mode == CGM_OP_GREEN || chi, CGM_ERROR_INVALID_VALUE, "null contrast");
  CGM_CHECK(chi_stride == 0 || chi_stride >= (long)c->N * c->N, CGM_ERROR_INVALID_VALUE,
            "chi_stride must be 0 (shared map) or >= N*N");
  const size_t need = sizeof(float2) * (size_t)c->N * c->M * (size_t)batch;
  CGM_CHECK(need <= kScratchBytes, CGM_ERROR_INVALID_VALUE,
            "cgm_matvec_batched_noalloc: %d vectors at N=%d need %zu MiB of FFT scratch, more than its %zu MiB "
            "(at most %zu vectors at this N)",
            batch, c->N, need >> 20, kScratchBytes >> 20, kScratchBytes / (sizeof(float2) * (size_t)c->N * c->M));
  void* p = nullptr;
  if (c->device >= 0 && c->device < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch_ptr[c->device];
  }
  CGM_CHECK(p != nullptr, CGM_ERROR_INVALID_VALUE, "no scratch address for device %d (context not created there)",
            c->device);
  CGM_HIP(hipSetDevice(c->device));
  float2* Tm = static_cast<float2*>(p);
  const long n = (long)c->N * c->N;
  BatchStrides bs;
  bs.x = n;
  bs.chi = chi_stride;
  bs.y = n;
  bs.T = (long)c->N * c->M;
  int rc = launch_pass1<float>(c->plan, c->stream, (const float2*)x, nullptr, chi, nullptr, mode, Tm, c->tw32,
                               nullptr, batch, bs);
  if (!rc) rc = launch_pass2<float, float2>(c->plan, c->stream, Tm, (const float2*)gamma, c->tw32, nullptr, batch, bs);
  if (!rc)
    rc = launch_pass3<float>(c->plan, c->stream, Tm, (const float2*)x, chi, nullptr, (float2*)y, mode, c->tw32,
                             nullptr, batch, bs);
  return rc;
}
