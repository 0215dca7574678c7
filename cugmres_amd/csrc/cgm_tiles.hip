// Tile (block-Jacobi) right preconditioner: exact inverses of the operator
// restricted to square tiles of the grid.
//
// The reference has no preconditioner (B/CUDA_GMRES/GMRES.cuh:22-446 runs
// plain GMRES).  On the high-contrast maps of the benchmark workloads
// (eps in {1, 12.1}, h = lambda/20) GMRES needs a number of steps that grows
// about as N^1.8 (335 at N = 64, 1,155 at N = 128 unrestarted), because every
// high-permittivity block is a near-resonant scatterer.  Inverting the
// operator exactly on T x T tiles removes those local resonances.
//
// On tile t (cells (a T + r, b T + c), 0 <= r, c < T, row-major inside the
// tile) the local operator is
//     A_t = I - G_t diag(chi_t),   G_t[p, q] = g[|p_r - q_r|, |p_c - q_c|],
// with g the N x N generator block of the operator (the block whose circulant
// embedding gives the kernel spectrum).  Factorisation, once per map: every
// A_t is formed in FP32 complex (column-major, T^2 x T^2) and inverted in
// place by a batched LU with partial pivoting (rocSOLVER getrf + getri).
// Application, once per Krylov step: y_t = A_t^{-1} x_t, one work-group per
// (tile, system), one thread per output cell streaming its row of the
// inverse column by column (coalesced: the inverse is column-major), the
// tile's input staged in LDS.  A map shared by several right-hand sides is
// factored once, and each work-group then applies its tile's inverse to all
// of them while it streams it.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <vector>

#include "cgm_internal.h"

using namespace cgm;

struct cgm_tiles {
  cgm_context* ctx = nullptr;
  int device = 0;
  int N = 0, T = 0, B = 0;  // grid, tile side, systems (maps) the workspace holds
  int nt = 0;               // tiles per grid side
  int M2 = 0;               // unknowns per tile (T * T)
  float2* K = nullptr;      // [B][nt * nt][M2 * M2] inverses, column-major
  int* ipiv = nullptr;      // [B][nt * nt][M2]
  int* info = nullptr;      // [B][nt * nt]
  int maps = 0;             // maps factored by the last cgm_tiles_factor (1: shared)
  rocblas_handle h = nullptr;
};

namespace {

// A_t of tile blockIdx.x of map blockIdx.y, column-major into K.
__global__ void __launch_bounds__(256)
k_tile_build(const double2* __restrict__ g, const float* __restrict__ chi, long chi_stride, float2* __restrict__ K,
             int N, int T) {
  const int nt = N / T;
  const int M2 = T * T;
  const int t = blockIdx.x;
  const int a = t / nt, b = t % nt;
  chi += (long)blockIdx.y * chi_stride;
  float2* A = K + ((long)blockIdx.y * nt * nt + t) * (long)M2 * M2;
  for (int e = threadIdx.x; e < M2 * M2; e += blockDim.x) {
    const int p = e % M2, q = e / M2;  // row p, column q
    const int pr = p / T, pc = p % T, qr = q / T, qc = q % T;
    const int dr = pr > qr ? pr - qr : qr - pr;
    const int dc = pc > qc ? pc - qc : qc - pc;
    const double2 gv = g[(long)dr * N + dc];
    const double cq = chi[(long)(a * T + qr) * N + b * T + qc];
    A[e] = make_float2((float)((p == q ? 1.0 : 0.0) - gv.x * cq), (float)(-gv.y * cq));
  }
}

// y_t = K_t x_t for NV vectors that share the tile's inverse (one map) or
// one vector per map (NV = 1).  blockIdx.x = tile, blockIdx.y = group of NV
// systems; system s reads io.v + s sV, scaled by 1 / *(io.nrm + s sNrm bytes)
// when io.nrm is given (the scaled input is written to io.vw + s sV when
// io.vw is given), writes io.z + s sZ, and is skipped once
// *(done + s sd bytes) is set.
template <int M2, int NV>
__global__ void __launch_bounds__(M2)
k_tile_apply(const float2* __restrict__ K, long sK, FnoIO io, const int* __restrict__ done, long sd, int N, int T,
             int nsys) {
  constexpr int kU = 8;  // columns of the inverse loaded ahead
  __shared__ float2 xs[NV][M2];
  const int nt = N / T;
  const int t = blockIdx.x;
  const int a = t / nt, b = t % nt;
  const int s0 = blockIdx.y * NV;
  const int r = threadIdx.x;
  const long cell = (long)(a * T + r / T) * N + b * T + r % T;
  bool live[NV];
  bool any = false;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int s = s0 + v;
    live[v] = s < nsys && !(done && *boff(done, (long)s * sd));
    any |= live[v];
    float2 xv = make_float2(0.f, 0.f);
    if (live[v]) {
      xv = io.v[(long)s * io.sV + cell];
      if (io.nrm) {
        const double d = *boff(io.nrm, (long)s * io.sNrm);
        const float sc = d > 0.0 ? (float)(1.0 / d) : 0.f;
        xv.x *= sc;
        xv.y *= sc;
        if (io.vw) io.vw[(long)s * io.sV + cell] = xv;
      }
    }
    xs[v][r] = xv;
  }
  if (!any) return;  // uniform over the work-group
  __syncthreads();
  const float2* Kt = K + (long)(NV == 1 ? s0 : 0) * sK + (long)t * M2 * M2 + r;
  float2 acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = make_float2(0.f, 0.f);
  for (int q0 = 0; q0 < M2; q0 += kU) {
    float2 k[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) k[u] = Kt[(long)(q0 + u) * M2];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const float2 xv = xs[v][q0 + u];
        acc[v].x += k[u].x * xv.x - k[u].y * xv.y;
        acc[v].y += k[u].x * xv.y + k[u].y * xv.x;
      }
    }
  }
#pragma unroll
  for (int v = 0; v < NV; ++v)
    if (live[v]) io.z[(long)(s0 + v) * io.sZ + cell] = acc[v];
}

template <int M2>
int launch_apply_m(const cgm_tiles* tl, hipStream_t st, const FnoIO& io, const int* done, long sd, int nsys) {
  const long sK = (long)tl->nt * tl->nt * M2 * M2;
  const int ntiles = tl->nt * tl->nt;
  if (tl->maps == 1 && nsys > 1) {
    constexpr int NV = 4;
    const int groups = (nsys + NV - 1) / NV;
    hipLaunchKernelGGL((k_tile_apply<M2, NV>), dim3(ntiles, groups), dim3(M2), 0, st, tl->K, 0L, io, done, sd,
                       tl->N, tl->T, nsys);
  } else {
    hipLaunchKernelGGL((k_tile_apply<M2, 1>), dim3(ntiles, nsys), dim3(M2), 0, st, tl->K, sK, io, done, sd, tl->N,
                       tl->T, nsys);
  }
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

}  // namespace

namespace cgm {

cgm_context* tiles_context(const cgm_tiles* t) { return t->ctx; }
int tiles_maps(const cgm_tiles* t) { return t->maps; }

int tiles_apply_io(const cgm_tiles* tl, hipStream_t st, const FnoIO& io, int nsys, const int* done, long sd) {
  if (tl->maps < 1) {
    cgm_set_last_error("the tile preconditioner has not been factored");
    return CGM_ERROR_INVALID_VALUE;
  }
  if (tl->maps > 1 && nsys > tl->maps) {
    cgm_set_last_error("%d systems but %d factored maps", nsys, tl->maps);
    return CGM_ERROR_INVALID_VALUE;
  }
  switch (tl->M2) {
    case 64: return launch_apply_m<64>(tl, st, io, done, sd, nsys);
    case 256: return launch_apply_m<256>(tl, st, io, done, sd, nsys);
    case 1024: return launch_apply_m<1024>(tl, st, io, done, sd, nsys);
    default:
      cgm_set_last_error("tile side %d unsupported", tl->T);
      return CGM_ERROR_UNSUPPORTED;
  }
}

}  // namespace cgm

extern "C" int cgm_tiles_create(cgm_ctx_t c, int T, int batch, cgm_tiles_t* out) {
  CGM_CHECK(c && out, CGM_ERROR_INVALID_VALUE, "null argument");
  *out = nullptr;
  CGM_CHECK(T == 8 || T == 16 || T == 32, CGM_ERROR_UNSUPPORTED, "tile side %d: 8, 16 or 32 supported", T);
  CGM_CHECK(c->N % T == 0, CGM_ERROR_UNSUPPORTED, "N=%d is not a multiple of the tile side %d", c->N, T);
  CGM_CHECK(batch >= 1, CGM_ERROR_INVALID_VALUE, "batch %d < 1", batch);
  CGM_HIP(hipSetDevice(c->device));
  cgm_tiles* t = new cgm_tiles();
  t->ctx = c;
  t->device = c->device;
  t->N = c->N;
  t->T = T;
  t->B = batch;
  t->nt = c->N / T;
  t->M2 = T * T;
  const size_t ntiles = (size_t)t->nt * t->nt;
  const size_t M2 = (size_t)t->M2;
  if (hipMalloc((void**)&t->K, sizeof(float2) * M2 * M2 * ntiles * batch) != hipSuccess ||
      hipMalloc((void**)&t->ipiv, sizeof(int) * M2 * ntiles * batch) != hipSuccess ||
      hipMalloc((void**)&t->info, sizeof(int) * ntiles * batch) != hipSuccess ||
      rocblas_create_handle(&t->h) != rocblas_status_success) {
    cgm_tiles_destroy(t);
    cgm_set_last_error("cgm_tiles_create: allocation failed (%zu MiB of tile inverses)",
                       sizeof(float2) * M2 * M2 * ntiles * batch >> 20);
    return CGM_ERROR_ALLOC;
  }
  *out = t;
  return CGM_SUCCESS;
}

extern "C" int cgm_tiles_destroy(cgm_tiles_t t) {
  if (!t) return CGM_SUCCESS;
  (void)hipSetDevice(t->device);
  void* ptrs[] = {t->K, t->ipiv, t->info};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (t->h) (void)rocblas_destroy_handle(t->h);
  delete t;
  return CGM_SUCCESS;
}

extern "C" int cgm_tiles_memory(cgm_tiles_t t, size_t* bytes) {
  CGM_CHECK(t && bytes, CGM_ERROR_INVALID_VALUE, "null argument");
  const size_t ntiles = (size_t)t->nt * t->nt;
  *bytes = (sizeof(float2) * t->M2 * t->M2 + sizeof(int) * (t->M2 + 1)) * ntiles * t->B;
  return CGM_SUCCESS;
}

extern "C" int cgm_tiles_factor(cgm_tiles_t t, const cgm_complex128* block, const float* chi, long chi_stride,
                                int maps) {
  CGM_CHECK(t && block && chi, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(maps >= 1 && maps <= t->B, CGM_ERROR_INVALID_VALUE, "maps %d outside [1, %d]", maps, t->B);
  CGM_CHECK(maps == 1 || chi_stride >= (long)t->N * t->N, CGM_ERROR_INVALID_VALUE,
            "chi_stride must be >= N*N for several maps");
  cgm_context* c = t->ctx;
  CGM_HIP(hipSetDevice(c->device));
  const int ntiles = t->nt * t->nt;
  const int M2 = t->M2;
  t->maps = 0;
  hipLaunchKernelGGL(k_tile_build, dim3(ntiles, maps), dim3(256), 0, c->stream, (const double2*)block, chi,
                     chi_stride, t->K, t->N, t->T);
  CGM_HIP(hipGetLastError());
  CGM_CHECK(rocblas_set_stream(t->h, c->stream) == rocblas_status_success, CGM_ERROR_DEVICE,
            "rocblas_set_stream failed");
  const rocblas_stride sA = (rocblas_stride)M2 * M2;
  const int nb = ntiles * maps;
  rocblas_status s = rocsolver_cgetrf_strided_batched(t->h, M2, M2, (rocblas_float_complex*)t->K, M2, sA, t->ipiv,
                                                      M2, t->info, nb);
  CGM_CHECK(s == rocblas_status_success, CGM_ERROR_DEVICE, "rocsolver getrf: status %d", (int)s);
  std::vector<int> info(nb);
  CGM_HIP(hipMemcpyAsync(info.data(), t->info, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream));
  CGM_HIP(hipStreamSynchronize(c->stream));
  for (int k = 0; k < nb; ++k)
    CGM_CHECK(info[k] == 0, CGM_ERROR_NONFINITE, "tile %d of map %d is singular (pivot %d)", k % ntiles,
              k / ntiles, info[k]);
  s = rocsolver_cgetri_strided_batched(t->h, M2, (rocblas_float_complex*)t->K, M2, sA, t->ipiv, M2, t->info, nb);
  CGM_CHECK(s == rocblas_status_success, CGM_ERROR_DEVICE, "rocsolver getri: status %d", (int)s);
  CGM_HIP(hipStreamSynchronize(c->stream));
  t->maps = maps;
  return CGM_SUCCESS;
}

extern "C" int cgm_tiles_apply(cgm_tiles_t t, const cgm_complex64* x, cgm_complex64* y, int batch) {
  CGM_CHECK(t && x && y, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(t->maps >= 1, CGM_ERROR_INVALID_VALUE, "cgm_tiles_factor has not run");
  CGM_CHECK(batch >= 1 && (t->maps == 1 || batch <= t->maps), CGM_ERROR_INVALID_VALUE,
            "batch %d: %d maps were factored", batch, t->maps);
  CGM_CHECK(x != y, CGM_ERROR_INVALID_VALUE, "x and y must not alias");
  CGM_HIP(hipSetDevice(t->device));
  FnoIO io;
  io.v = (const float2*)x;
  io.sV = (long)t->N * t->N;
  io.z = (float2*)y;
  io.sZ = io.sV;
  return tiles_apply_io(t, t->ctx->stream, io, batch, nullptr, 0);
}
