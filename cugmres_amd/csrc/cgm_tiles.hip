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
//
// 64 x 64 tiles (4,096 unknowns each) store their inverses in FP16 complex
// (half the bytes per step), each tile scaled by its largest magnitude so no
// entry leaves FP16's range; the scale is applied to the FP32 sums.  A cell q
// with chi_q = 0 makes column q of A_t the unit vector e_q, so column q of the
// inverse is e_q as well: only the columns of cells with chi != 0 are stored
// (packed in order) and streamed, and the others add x_q to y_q.  On the
// benchmark maps that is about three columns in eight.  The application
// splits every tile's rows over 8 work-groups.  On the benchmark
// maps at N = 512 they halve the Krylov steps of 32 x 32 tiles (2,391 against
// 4,573 on seed 0), and FP16 storage leaves the step count unchanged.
#include <hip/hip_fp16.h>
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
  __half2* K16 = nullptr;   // FP16 copy (T = 64): [B][nt * nt][column slot][M2] of K / kscale,
                            // only the columns of cells with chi != 0 (the others are unit vectors)
  float* kscale = nullptr;  // [B][nt * nt] largest |K| entry of each tile's kept columns (T = 64)
  int* ncol = nullptr;      // [B][nt * nt] kept columns, padded to a multiple of 16 (T = 64)
  unsigned short* cols = nullptr;  // [B][nt * nt][M2] column of each slot (T = 64)
  unsigned int* zmask = nullptr;   // [B][nt * nt][M2 / 32] cells with chi = 0 (T = 64)
  std::vector<double> kept;        // per map: kept columns over all tiles (host; bytes per application)
  int* ipiv = nullptr;      // [B][nt * nt][M2]
  int* info = nullptr;      // [B][nt * nt]
  int maps = 0;             // maps factored by the last cgm_tiles_factor (1: shared)
  rocblas_handle h = nullptr;
};

namespace {

constexpr int kBuildChunk = 65536;  // matrix entries per work-group of k_tile_build

// A_t of map blockIdx.y, column-major into K: work-group blockIdx.x covers
// chunk (blockIdx.x % chunks) of tile (blockIdx.x / chunks).
__global__ void __launch_bounds__(256)
k_tile_build(const double2* __restrict__ g, const float* __restrict__ chi, long chi_stride, float2* __restrict__ K,
             int N, int T, int chunks) {
  const int nt = N / T;
  const int M2 = T * T;
  const int t = blockIdx.x / chunks;
  const long e0 = (long)(blockIdx.x % chunks) * kBuildChunk;
  const long e1 = e0 + kBuildChunk < (long)M2 * M2 ? e0 + kBuildChunk : (long)M2 * M2;
  const int a = t / nt, b = t % nt;
  chi += (long)blockIdx.y * chi_stride;
  float2* A = K + ((long)blockIdx.y * nt * nt + t) * (long)M2 * M2;
  for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
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

constexpr int kTileH = 4096;           // unknowns of a 64 x 64 tile
constexpr int kRowsH = 512;            // rows per work-group of the application
constexpr int kPartsH = kTileH / kRowsH;
constexpr int kColPad = 16;            // kept columns are padded to a multiple of this

constexpr unsigned short kPadCol = 0xFFFF;  // column index of a padding slot

// Kept columns of every 64 x 64 tile of map blockIdx.y (tile blockIdx.x):
// the cells with chi != 0 in order, their count padded to kColPad (padding
// slots hold kPadCol and zero columns), and the chi = 0 bitmask.
__global__ void __launch_bounds__(1024)
k_tile_cols(const float* __restrict__ chi, long chi_stride, int N, int* __restrict__ ncol,
            unsigned short* __restrict__ cols, unsigned int* __restrict__ zmask) {
  constexpr int T = 64;
  const int nt = N / T;
  const int t = blockIdx.x, a = t / nt, b = t % nt;
  const long tb = (long)blockIdx.y * nt * nt + t;
  chi += (long)blockIdx.y * chi_stride;
  __shared__ unsigned int zm[kTileH / 32];
  __shared__ int wsum[16];
  for (int w = threadIdx.x; w < kTileH / 32; w += blockDim.x) zm[w] = 0u;
  __syncthreads();
  bool nz[4];
  int cnt = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int c = 4 * threadIdx.x + i;
    nz[i] = chi[(long)(a * T + c / T) * N + b * T + c % T] != 0.f;
    cnt += nz[i];
    if (!nz[i]) atomicOr(&zm[c >> 5], 1u << (c & 31));
  }
  // exclusive scan of the counts over the work-group (waves of 64)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int base = 0;
  for (int w = 0; w < wave; ++w) base += wsum[w];
  int slot = base + inc - cnt;
  unsigned short* ct = cols + tb * kTileH;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (nz[i]) ct[slot++] = (unsigned short)(4 * threadIdx.x + i);
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) total += wsum[w];
    const int padded = (total + kColPad - 1) / kColPad * kColPad;
    for (int e = total; e < padded; ++e) ct[e] = kPadCol;
    ncol[tb] = padded;
  }
  for (int w = threadIdx.x; w < kTileH / 32; w += blockDim.x) zmask[tb * (kTileH / 32) + w] = zm[w];
}

// Largest |K| entry of each tile's kept columns (one work-group per tile of one map).
__global__ void __launch_bounds__(1024) k_tile_maxabs(const float2* __restrict__ K, const int* __restrict__ ncol,
                                                      const unsigned short* __restrict__ cols, float* __restrict__ sc) {
  const long tb = blockIdx.x;
  const float2* Kt = K + tb * (long)kTileH * kTileH;
  const int nc = ncol[tb];
  float mx = 0.f;
  for (long e = threadIdx.x; e < (long)nc * kTileH; e += blockDim.x) {
    const int s = (int)(e / kTileH), p = (int)(e % kTileH);
    const int q = cols[tb * kTileH + s];
    if (q == kPadCol) continue;
    const float2 v = Kt[(long)q * kTileH + p];
    mx = fmaxf(mx, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  __shared__ float red[16];
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmaxf(m, red[w]);
    sc[tb] = m > 0.f ? m : 1.f;
  }
}

// K16 slot s of tile tb = column cols[s] of K / scale (zero for padding slots).
__global__ void __launch_bounds__(256) k_tile_to_half(const float2* __restrict__ K, __half2* __restrict__ K16,
                                                      const float* __restrict__ sc, const int* __restrict__ ncol,
                                                      const unsigned short* __restrict__ cols, long ntile_total) {
  const long m22 = (long)kTileH * kTileH;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < m22 * ntile_total; e += (long)gridDim.x * blockDim.x) {
    const long tb = e / m22;
    const long w = e % m22;
    const int s = (int)(w / kTileH), p = (int)(w % kTileH);
    if (s >= ncol[tb]) continue;
    const int q = cols[tb * kTileH + s];
    const float r = 1.f / sc[tb];
    const float2 v = q == kPadCol ? make_float2(0.f, 0.f) : K[tb * m22 + (long)q * kTileH + p];
    K16[e] = __floats2half2_rn(v.x * r, v.y * r);
  }
}

// y_t = scale_t K16_t x_t + (x_q on the cells with chi = 0) for 64 x 64
// tiles: work-group (tile, part) of system blockIdx.y computes the kRowsH rows
// [part kRowsH, (part+1) kRowsH) of its tile, two consecutive rows per thread
// (one 8-byte load of a packed FP16 column per kept column), FP32 sums; the
// tile's input and its column list are staged in LDS by every part (part 0
// also writes the scaled input back to io.vw).  Same io / done conventions as
// k_tile_apply with NV = 1.
__global__ void __launch_bounds__(kRowsH / 2)
k_tile_apply_h(const __half2* __restrict__ K16, const float* __restrict__ ksc, const int* __restrict__ ncol,
               const unsigned short* __restrict__ cols, const unsigned int* __restrict__ zmask, int maps1, FnoIO io,
               const int* __restrict__ done, long sd, int N) {
  constexpr int T = 64;
  constexpr int kU = 16;  // columns loaded ahead (kColPad divides every padded count)
  __shared__ float2 xs[kTileH];
  __shared__ unsigned short cs[kTileH];
  const int nt = N / T;
  const int t = blockIdx.x / kPartsH, part = blockIdx.x % kPartsH;
  const int a = t / nt, b = t % nt;
  const int s = blockIdx.y;
  if (done && *boff(done, (long)s * sd)) return;
  const long tb = (long)(maps1 ? 0 : s) * nt * nt + t;
  float sc = 1.f;
  if (io.nrm) {
    const double d = *boff(io.nrm, (long)s * io.sNrm);
    sc = d > 0.0 ? (float)(1.0 / d) : 0.f;
  }
  const int nc = ncol[tb];
  for (int r = threadIdx.x; r < kTileH; r += blockDim.x) {
    const long cell = (long)(a * T + r / T) * N + b * T + r % T;
    float2 xv = io.v[(long)s * io.sV + cell];
    xv.x *= sc;
    xv.y *= sc;
    if (io.nrm && io.vw && part == 0) io.vw[(long)s * io.sV + cell] = xv;
    xs[r] = xv;
    if (r < nc) {
      const unsigned short c = cols[tb * kTileH + r];
      cs[r] = c == kPadCol ? 0 : c;  // a padding slot's column is zero: any input works
    }
  }
  __syncthreads();
  const int p = part * kRowsH + 2 * threadIdx.x;  // rows p, p + 1
  const uint2* Kt = reinterpret_cast<const uint2*>(K16 + tb * (long)kTileH * kTileH + p);
  float ax = 0.f, ay = 0.f, bx = 0.f, by = 0.f;
  for (int q0 = 0; q0 < nc; q0 += kU) {
    uint2 k[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      // the inverses stream once per application: non-temporal loads
      typedef unsigned int u2v_t __attribute__((ext_vector_type(2)));
      const u2v_t w = __builtin_nontemporal_load(reinterpret_cast<const u2v_t*>(Kt + (long)(q0 + u) * (kTileH / 2)));
      k[u] = make_uint2(w.x, w.y);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const float2 xv = xs[cs[q0 + u]];
      const float2 k0 = __half22float2(__builtin_bit_cast(__half2, k[u].x));
      const float2 k1 = __half22float2(__builtin_bit_cast(__half2, k[u].y));
      ax += k0.x * xv.x - k0.y * xv.y;
      ay += k0.x * xv.y + k0.y * xv.x;
      bx += k1.x * xv.x - k1.y * xv.y;
      by += k1.x * xv.y + k1.y * xv.x;
    }
  }
  const float m = ksc[tb];
  const unsigned int zw = zmask[tb * (kTileH / 32) + (p >> 5)];
  const bool z0 = (zw >> (p & 31)) & 1u, z1 = (zw >> ((p + 1) & 31)) & 1u;   // p even: same word
  float2 y0 = make_float2(ax * m, ay * m), y1 = make_float2(bx * m, by * m);
  if (z0) y0 = cadd(y0, xs[p]);
  if (z1) y1 = cadd(y1, xs[p + 1]);
  const long c0 = (long)(a * T + p / T) * N + b * T + p % T;   // p and p + 1 share a tile row (p even)
  io.z[(long)s * io.sZ + c0] = y0;
  io.z[(long)s * io.sZ + c0 + 1] = y1;
}

int launch_apply_h(const cgm_tiles* tl, hipStream_t st, const FnoIO& io, const int* done, long sd, int nsys) {
  const int ntiles = tl->nt * tl->nt;
  hipLaunchKernelGGL(k_tile_apply_h, dim3(ntiles * kPartsH, nsys), dim3(kRowsH / 2), 0, st, tl->K16, tl->kscale,
                     tl->ncol, tl->cols, tl->zmask, tl->maps == 1 ? 1 : 0, io, done, sd, tl->N);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
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

// Algorithmic bytes of one application to nsys systems: the columns streamed
// (every column of every tile in FP32, once per system or once per group of
// four right-hand sides of one map; the kept FP16 columns of the map each
// system uses for 64 x 64 tiles) plus v in and z out.
double apply_bytes(const cgm_tiles* tl, int nsys) {
  const double per_entry = tl->K16 ? 4.0 : 8.0;
  double cols = 0.0;
  if (tl->maps == 1)
    cols = tl->kept[0] * ((nsys > 1 && tl->M2 < 4096) ? (nsys + 3) / 4 : nsys);
  else
    for (int k = 0; k < nsys; ++k) cols += tl->kept[k];
  return cols * tl->M2 * per_entry + 16.0 * tl->N * tl->N * nsys;
}

int tiles_apply_io(const cgm_tiles* tl, hipStream_t st, const FnoIO& io, int nsys, const int* done, long sd) {
  if (tl->maps < 1) {
    cgm_set_last_error("the tile preconditioner has not been factored");
    return CGM_ERROR_INVALID_VALUE;
  }
  if (tl->maps > 1 && nsys > tl->maps) {
    cgm_set_last_error("%d systems but %d factored maps", nsys, tl->maps);
    return CGM_ERROR_INVALID_VALUE;
  }
  // live timing inside a long cycle (cgm_longrec_timing): events around the
  // launch and its algorithmic bytes
  LrTiming* tm = tl->ctx->lrt.on ? &tl->ctx->lrt : nullptr;
  const int sl = tm ? tm->cur : -1;
  constexpr int KK = LrTiming::kKernels;
  if (sl >= 0) CGM_HIP(hipEventRecord(tm->ev[2 * KK * sl + 4], st));
  int rc;
  switch (tl->M2) {
    case 64: rc = launch_apply_m<64>(tl, st, io, done, sd, nsys); break;
    case 256: rc = launch_apply_m<256>(tl, st, io, done, sd, nsys); break;
    case 1024: rc = launch_apply_m<1024>(tl, st, io, done, sd, nsys); break;
    case 4096: rc = launch_apply_h(tl, st, io, done, sd, nsys); break;
    default:
      cgm_set_last_error("tile side %d unsupported", tl->T);
      return CGM_ERROR_UNSUPPORTED;
  }
  if (sl >= 0 && !rc) {
    CGM_HIP(hipEventRecord(tm->ev[2 * KK * sl + 5], st));
    tm->slot_bytes[2][sl] = apply_bytes(tl, nsys);
    tm->slot_has[2][sl] = true;
  }
  return rc;
}

}  // namespace cgm

extern "C" int cgm_tiles_create(cgm_ctx_t c, int T, int batch, cgm_tiles_t* out) {
  CGM_CHECK(c && out, CGM_ERROR_INVALID_VALUE, "null argument");
  *out = nullptr;
  CGM_CHECK(T == 8 || T == 16 || T == 32 || T == 64, CGM_ERROR_UNSUPPORTED, "tile side %d: 8, 16, 32 or 64 supported",
            T);
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
  const bool half = T == 64;
  if (hipMalloc((void**)&t->K, sizeof(float2) * M2 * M2 * ntiles * batch) != hipSuccess ||
      hipMalloc((void**)&t->ipiv, sizeof(int) * M2 * ntiles * batch) != hipSuccess ||
      hipMalloc((void**)&t->info, sizeof(int) * ntiles * batch) != hipSuccess ||
      (half && hipMalloc((void**)&t->K16, sizeof(__half2) * M2 * M2 * ntiles * batch) != hipSuccess) ||
      (half && hipMalloc((void**)&t->kscale, sizeof(float) * ntiles * batch) != hipSuccess) ||
      (half && hipMalloc((void**)&t->ncol, sizeof(int) * ntiles * batch) != hipSuccess) ||
      (half && hipMalloc((void**)&t->cols, sizeof(unsigned short) * M2 * ntiles * batch) != hipSuccess) ||
      (half && hipMalloc((void**)&t->zmask, sizeof(unsigned int) * (M2 / 32) * ntiles * batch) != hipSuccess) ||
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
  void* ptrs[] = {t->K, t->ipiv, t->info, t->K16, t->kscale, t->ncol, t->cols, t->zmask};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (t->h) (void)rocblas_destroy_handle(t->h);
  delete t;
  return CGM_SUCCESS;
}

extern "C" int cgm_tiles_memory(cgm_tiles_t t, size_t* bytes) {
  CGM_CHECK(t && bytes, CGM_ERROR_INVALID_VALUE, "null argument");
  const size_t ntiles = (size_t)t->nt * t->nt;
  const size_t k16 = t->K16 ? sizeof(__half2) * t->M2 * t->M2 + sizeof(float) + sizeof(int) +
                                  sizeof(unsigned short) * t->M2 + sizeof(unsigned int) * (t->M2 / 32)
                            : 0;
  *bytes = (sizeof(float2) * t->M2 * t->M2 + sizeof(int) * (t->M2 + 1) + k16) * ntiles * t->B;
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
  const int chunks = (int)(((long)M2 * M2 + kBuildChunk - 1) / kBuildChunk);
  hipLaunchKernelGGL(k_tile_build, dim3(ntiles * chunks, maps), dim3(256), 0, c->stream, (const double2*)block, chi,
                     chi_stride, t->K, t->N, t->T, chunks);
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
  t->kept.assign(maps, 0.0);
  if (t->K16) {
    hipLaunchKernelGGL(k_tile_cols, dim3(ntiles, maps), dim3(1024), 0, c->stream, chi, chi_stride, t->N, t->ncol,
                       t->cols, t->zmask);
    hipLaunchKernelGGL(k_tile_maxabs, dim3(nb), dim3(1024), 0, c->stream, t->K, t->ncol, t->cols, t->kscale);
    hipLaunchKernelGGL(k_tile_to_half, dim3(4096), dim3(256), 0, c->stream, t->K, t->K16, t->kscale, t->ncol,
                       t->cols, (long)nb);
    CGM_HIP(hipGetLastError());
    std::vector<int> nc(nb);
    CGM_HIP(hipMemcpyAsync(nc.data(), t->ncol, sizeof(int) * nb, hipMemcpyDeviceToHost, c->stream));
    CGM_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < nb; ++k) t->kept[k / ntiles] += nc[k];
  } else {
    for (int m = 0; m < maps; ++m) t->kept[m] = (double)ntiles * M2;
  }
  CGM_HIP(hipStreamSynchronize(c->stream));
  t->maps = maps;
  return CGM_SUCCESS;
}

extern "C" int cgm_tiles_apply_bytes(cgm_tiles_t t, int batch, double* bytes) {
  CGM_CHECK(t && bytes, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(t->maps >= 1, CGM_ERROR_INVALID_VALUE, "cgm_tiles_factor has not run");
  CGM_CHECK(batch >= 1 && (t->maps == 1 || batch <= t->maps), CGM_ERROR_INVALID_VALUE,
            "batch %d: %d maps were factored", batch, t->maps);
  *bytes = apply_bytes(t, batch);
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

// ---------------------------------------------------------------------------
// diagnostics: the application alone on synthetic inverses
// ---------------------------------------------------------------------------
namespace {
// Synthetic inverses for one map; 64 x 64 tiles keep every column.
__global__ void k_tile_fill(float2* __restrict__ K, __half2* __restrict__ K16, float* __restrict__ sc,
                            int* __restrict__ ncol, unsigned short* __restrict__ cols, unsigned int* __restrict__ zmask,
                            int M2, long total, long ntiles) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const float v = 1e-3f * (float)(e % 7);
    if (K16)
      K16[e] = __floats2half2_rn(v, -v);
    else
      K[e] = make_float2(v, -v);
    if (sc && e < ntiles) {
      sc[e] = 1.f;
      ncol[e] = M2;
    }
    if (cols && e < ntiles * M2) cols[e] = (unsigned short)(e % M2);
    if (zmask && e < ntiles * (M2 / 32)) zmask[e] = 0u;
  }
}
}  // namespace

extern "C" int cgm_profile_tiles(cgm_tiles_t t, int reps, float* ms) {
  CGM_CHECK(t && ms, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(reps >= 1, CGM_ERROR_INVALID_VALUE, "reps < 1");
  cgm_context* c = t->ctx;
  CGM_HIP(hipSetDevice(c->device));
  const long ntiles = (long)t->nt * t->nt;
  const long total = (long)t->M2 * t->M2 * ntiles;  // one map
  hipLaunchKernelGGL(k_tile_fill, dim3(4096), dim3(256), 0, c->stream, t->K, t->K16, t->kscale, t->ncol, t->cols,
                     t->zmask, t->M2, total, ntiles);
  CGM_HIP(hipGetLastError());
  const long n = (long)t->N * t->N;
  float2 *x = nullptr, *y = nullptr;
  CGM_HIP(hipMalloc((void**)&x, sizeof(float2) * n));
  if (hipMalloc((void**)&y, sizeof(float2) * n) != hipSuccess) {
    (void)hipFree(x);
    cgm_set_last_error("cgm_profile_tiles: allocation failed");
    return CGM_ERROR_ALLOC;
  }
  int rc = CGM_SUCCESS;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipMemsetAsync(x, 0, sizeof(float2) * n, c->stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = CGM_ERROR_DEVICE;
  } else {
    t->maps = 1;
    t->kept.assign(1, (double)ntiles * t->M2);
    FnoIO io;
    io.v = x;
    io.sV = n;
    io.z = y;
    io.sZ = n;
    rc = tiles_apply_io(t, c->stream, io, 1, nullptr, 0);
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; !rc && r < reps; ++r) rc = tiles_apply_io(t, c->stream, io, 1, nullptr, 0);
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(ms, e0, e1);
    *ms /= reps;
    t->maps = 0;  // the synthetic inverses must be factored before use
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(x);
  (void)hipFree(y);
  return rc;
}

