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
// embedding gives the kernel spectrum).
//
// Tiles of side 8, 16, 32 (dense form): every A_t is formed in FP32 complex
// (column-major, T^2 x T^2) and inverted in place by a batched LU with
// partial pivoting (rocSOLVER getrf + getri), once per map.  Application, once
// per Krylov step: y_t = A_t^{-1} x_t, one work-group per (tile, system), one
// thread per output cell streaming its row of the inverse column by column
// (coalesced: the inverse is column-major), the tile's input staged in LDS.
// A map shared by several right-hand sides is factored once, and each
// work-group then applies its tile's inverse to all of them while it streams
// it.
//
// Tiles of side 64 and 128 (reduced form).  With S the cells of the tile
// where chi != 0 and D_S = diag(chi_S), the columns of A_t outside S are unit
// vectors, and
//     A_t^{-1} = I + G_t[:, S] W_t R_S,   W_t = (I - D_S G_SS)^{-1} D_S
// (R_S restricts to S).  Factorisation inverts only the |S| x |S| matrices
// D_S^{-1} - G_SS = D_S^{-1} (I - D_S G_SS), whose inverse is W_t (a symmetric
// block sweep with rocBLAS updates and an LDS pivot-block kernel), and keeps
// W_t in FP16 complex, scaled per tile by its largest magnitude.  Application:
// u_t = W_t x_t[S] (the only large stream: W_t is complex symmetric, since
// W_t = (D_S^{-1} - G_SS)^{-1}, and only the lower triangle of its 128 x 128
// blocks is stored and streamed), then G_t u_t for all tiles at once as a
// batch of T x T convolutions through the operator's own FFT passes at size
// 2T with the tile's spectrum, then y = x + G_t u_t.  On the benchmark maps
// 3/8 of the cells carry chi != 0, so a 128 x 128 tile streams about 4 % of
// the entries of its dense inverse, and the inversion work drops by about
// 20x; that makes 128 x 128 tiles affordable, which take 1,032 Krylov steps
// on seed 0 at N = 512 against 2,380 for 64 x 64 and 4,573 for 32 x 32.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <map>
#include <vector>

#include "cgm_internal.h"
#include "cgm_special.h"

using namespace cgm;

struct cgm_tiles {
  cgm_context* ctx = nullptr;
  int device = 0;
  int N = 0, T = 0, B = 0;  // grid, tile side, systems (maps) the workspace holds
  int nt = 0;               // tiles per grid side
  int M2 = 0;               // unknowns per tile (T * T)
  bool reduced = false;     // T >= 64: reduced form
  int maps = 0;             // maps factored by the last cgm_tiles_factor (1: shared)
  rocblas_handle h = nullptr;
  size_t bytes = 0;         // device memory held
  // dense form (T <= 32)
  float2* K = nullptr;       // [B][nt * nt][M2 * M2] inverses, column-major
  int* ipiv = nullptr;      // [B][nt * nt][M2]
  int* info = nullptr;      // [B][nt * nt]       (reduced form: singular-pivot flags)
  // reduced form (T >= 64)
  int* ncol = nullptr;             // [B][nt * nt] |S_t| padded to a multiple of kPadTo
  unsigned short* cols = nullptr;  // [B][nt * nt][M2] the cells of S_t in order, then kPadCol
  long* woff = nullptr;            // [B][nt * nt] offset of W_t in W16 (half2 elements)
  float* kscale = nullptr;         // [B][nt * nt] largest |W_t| entry
  int* ids = nullptr;              // [B][nt * nt] factor members (tiles with n_t > 0), by size
  int* eloc = nullptr;             // [B][nt * nt] index of each tile's first block in its map's block list
  int* btab = nullptr;             // block lists of the maps: tile << 16 | block row << 8 | block column
  size_t btab_cap = 0;
  int* bstart = nullptr;           // [B] first entry of each map's block list in btab
  int* bcount = nullptr;           // [B] blocks of each map
  int bcount_max = 0;              // largest block count over the factored maps
  int nbmax = 0;                   // largest number of block rows of a tile
  float2* Pw = nullptr;            // [sys][bcount_max][2][kBlkW] partial products (apply)
  size_t pw_cap = 0;
  long* koff = nullptr;            // [B][nt * nt] offset of each member's matrix in Kw
  int* nl = nullptr;               // [B][nt * nt] each member's matrix side (n_t padded to kBlk)
  __half2* W16 = nullptr;          // W_t / kscale_t in FP16: the kBlkW x kBlkW blocks (i, j <= i) of the
                                   // lower triangle, each column-major, row-major over the blocks
  size_t w16_cap = 0;              // elements
  float2* Kw = nullptr;            // FP32 matrices of the members (factor)
  size_t kw_cap = 0;               // elements
  float2* Pb = nullptr;            // [member][kBlk][kBlk] pivot blocks (factor)
  size_t pb_cap = 0;
  float2* Rw = nullptr;            // kBlk x n_max row block (factor)
  size_t rw_cap = 0;
  float2* Cw = nullptr;            // n_max x kBlk column block (factor)
  size_t cw_cap = 0;
  float2* U = nullptr;             // [sys][nt * nt][M2] u_t scattered on the tile (apply)
  float2* Y = nullptr;             // [sys][nt * nt][M2] G_t u_t (apply)
  float2* Tm = nullptr;            // [sys][nt * nt][T * 2T] FFT scratch (apply)
  int scratch_sys = 0;             // systems the apply scratch holds
  Plan P;                          // tile convolution: N' = T, M' = 2T
  float2* tw32 = nullptr;          // twiddles of size M'
  double2* tw64 = nullptr;
  double2* blkT = nullptr;         // T x T corner of the generator block
  float2* gq = nullptr;            // quarter spectrum of the tile convolution
  std::vector<double> kept;        // per map: inverse entries streamed per application
};

namespace {

constexpr int kBuildChunk = 65536;           // matrix entries per work-group of the build kernels
constexpr int kPadTo = 16;                   // n_t is padded to a multiple of this
constexpr unsigned short kPadCol = 0xFFFF;   // cell index of a padding slot
constexpr int kBlk = 128;                    // block of the Gauss-Jordan inversion (reduced form)
constexpr int kBlkW = 128;                   // block of the symmetric storage of W (apply)

template <typename P>
int grow(cgm_tiles* t, P** p, size_t* cap, size_t need) {
  if (need <= *cap) return CGM_SUCCESS;
  if (*p) {
    (void)hipFree(*p);
    t->bytes -= *cap * sizeof(P);
    *p = nullptr;
    *cap = 0;
  }
  if (hipMalloc((void**)p, need * sizeof(P)) != hipSuccess) {
    *p = nullptr;
    cgm_set_last_error("tile preconditioner: allocation of %zu MiB failed", need * sizeof(P) >> 20);
    return CGM_ERROR_ALLOC;
  }
  *cap = need;
  t->bytes += need * sizeof(P);
  return CGM_SUCCESS;
}

// ---------------------------------------------------------------------------
// dense form (T <= 32)
// ---------------------------------------------------------------------------

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
#pragma unroll 2
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

// ---------------------------------------------------------------------------
// reduced form (T = 64, 128): factorisation
// ---------------------------------------------------------------------------

// S_t of every tile of map blockIdx.y (tile blockIdx.x): the cells with
// chi != 0 in row-major order and their count padded to kPadTo (padding
// slots hold kPadCol).
template <int T>
__global__ void __launch_bounds__(1024)
k_tile_cols(const float* __restrict__ chi, long chi_stride, int N, int* __restrict__ ncol,
            unsigned short* __restrict__ cols) {
  constexpr int M2 = T * T;
  constexpr int CPT = M2 / 1024;  // cells per thread
  const int nt = N / T;
  const int t = blockIdx.x, a = t / nt, b = t % nt;
  const long tb = (long)blockIdx.y * nt * nt + t;
  chi += (long)blockIdx.y * chi_stride;
  __shared__ int wsum[16];
  bool nz[CPT];
  int cnt = 0;
#pragma unroll
  for (int i = 0; i < CPT; ++i) {
    const int c = CPT * threadIdx.x + i;
    nz[i] = chi[(long)(a * T + c / T) * N + b * T + c % T] != 0.f;
    cnt += nz[i];
  }
  // exclusive scan of the counts over the work-group (16 waves of 64)
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
  unsigned short* ct = cols + tb * M2;
#pragma unroll
  for (int i = 0; i < CPT; ++i)
    if (nz[i]) ct[slot++] = (unsigned short)(CPT * threadIdx.x + i);
  int total = 0;
  for (int w = 0; w < 16; ++w) total += wsum[w];
  const int padded = (total + kPadTo - 1) / kPadTo * kPadTo;
  if ((int)threadIdx.x < padded - total) ct[total + threadIdx.x] = kPadCol;
  if (threadIdx.x == 0) ncol[tb] = padded;
}

// Grid index of cell c (row-major inside the tile) of tile t.
__device__ inline long tile_cell(int t, int c, int nt, int T, int N) {
  const int a = t / nt, b = t % nt;
  return (long)(a * T + c / T) * N + b * T + c % T;
}

// A = D_S^{-1} - G_SS of member blockIdx.y (tile ids[.]) into Kw + koff[.],
// nl x nl column-major, the identity on padding slots and beyond n_t (A is
// complex symmetric: G is, D_S is real diagonal; W_t = A^{-1}); work-group
// blockIdx.x covers one chunk of the entries.
__global__ void __launch_bounds__(256)
k_tile_build_red(const double2* __restrict__ g, const float* __restrict__ chi, long chi_stride,
                 const int* __restrict__ ncol, const unsigned short* __restrict__ cols, const int* __restrict__ ids,
                 const long* __restrict__ koff, const int* __restrict__ nl, float2* __restrict__ Kw, int N, int T) {
  const int nt = N / T;
  const int i = blockIdx.y;
  const int n = nl[i];
  const long nn = (long)n * n;
  const long e0 = (long)blockIdx.x * kBuildChunk;
  if (e0 >= nn) return;
  const long e1 = e0 + kBuildChunk < nn ? e0 + kBuildChunk : nn;
  const int tb = ids[i];
  const int map = tb / (nt * nt), t = tb % (nt * nt);
  const int nk = ncol[tb];
  const unsigned short* ct = cols + (long)tb * T * T;
  chi += (long)map * chi_stride;
  float2* A = Kw + koff[i];
  for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int p = (int)(e % n), q = (int)(e / n);
    const int cp = p < nk ? ct[p] : kPadCol, cq = q < nk ? ct[q] : kPadCol;
    float2 v = make_float2(p == q ? 1.f : 0.f, 0.f);
    if (cp != kPadCol && cq != kPadCol) {
      const int pr = cp / T, pc = cp % T, qr = cq / T, qc = cq % T;
      const int dr = pr > qr ? pr - qr : qr - pr;
      const int dc = pc > qc ? pc - qc : qc - pc;
      const double2 gv = g[(long)dr * N + dc];
      const double d = p == q ? 1.0 / (double)chi[tile_cell(t, cp, nt, T, N)] : 0.0;
      v = make_float2((float)(d - gv.x), (float)(-gv.y));
    }
    A[e] = v;
  }
}

// Pb[i] = the symmetric kBlk x kBlk diagonal block at k0 of member i
// (column-major), mirrored from the lower triangle where the sweep keeps it.
__global__ void __launch_bounds__(256)
k_gj_gather(const float2* __restrict__ Kw, const long* __restrict__ koff, const int* __restrict__ nl, int k0,
            float2* __restrict__ Pb) {
  const int i = blockIdx.x;
  const int n = nl[i];
  const float2* M = Kw + koff[i];
  float2* P = Pb + (long)i * kBlk * kBlk;
  for (int e = threadIdx.x; e < kBlk * kBlk; e += blockDim.x) {
    const int r = e % kBlk, c = e / kBlk;
    const int hi = r > c ? r : c, lo = r > c ? c : r;
    P[e] = M[(long)(k0 + lo) * n + k0 + hi];
  }
}

// Column block C (n x kBlk, column-major) of pivot block k0 of one member's
// symmetric matrix M (lower triangle kept): rows i < k0 from the row block
// (M(k0 + c, i)), rows in the pivot block mirrored, rows i >= k0 + kBlk from
// the column block (M(i, k0 + c)).  32 x 32 tiles through LDS, so that the
// transposed part reads and writes along columns.
__global__ void __launch_bounds__(256)
k_sw_gather(const float2* __restrict__ M, int n, int k0, float2* __restrict__ C) {
  __shared__ float2 tile[32][33];
  const int i0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    // tile[a][b] = C(i0 + a, c0 + b), every read along a column of M
    if (i0 < k0) {  // C(i0 + k, c0 + tx) = M(k0 + c0 + tx, i0 + k): column i0 + k, contiguous in tx
      tile[k][tx] = M[(long)(i0 + k) * n + k0 + c0 + tx];
    } else {
      const int i = i0 + tx, c = k0 + c0 + k;  // C(i, c - k0) = M(max, min) of (i, c)
      const int hi = i > c ? i : c, lo = i > c ? c : i;
      tile[tx][k] = M[(long)lo * n + hi];
    }
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) C[(long)(c0 + k) * n + i0 + tx] = tile[tx][k];
}

// Writes back the swept column block: CP = C P (n x kBlk) goes to rows
// i >= k0 + kBlk of the column block and, transposed, to the row block for
// rows i < k0; the pivot block becomes -P (its lower triangle is used).
__global__ void __launch_bounds__(256)
k_sw_scatter(float2* __restrict__ M, int n, int k0, const float2* __restrict__ CP, const float2* __restrict__ P) {
  __shared__ float2 tile[32][33];
  const int i0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (i0 >= k0 && i0 < k0 + kBlk) {  // the pivot block: -P
    for (int k = ty; k < 32; k += 8) {
      const float2 v = P[(long)(c0 + k) * kBlk + i0 - k0 + tx];
      M[(long)(k0 + c0 + k) * n + i0 + tx] = make_float2(-v.x, -v.y);
    }
    return;
  }
  if (i0 >= k0) {  // below: column block, straight copy along columns
    for (int k = ty; k < 32; k += 8) M[(long)(k0 + c0 + k) * n + i0 + tx] = CP[(long)(c0 + k) * n + i0 + tx];
    return;
  }
  for (int k = ty; k < 32; k += 8) tile[k][tx] = CP[(long)(c0 + k) * n + i0 + tx];  // CP(i0 + tx, c0 + k)
  __syncthreads();
  // M(k0 + c0 + tx, i0 + k) = CP(i0 + k, c0 + tx): column i0 + k of M, contiguous in tx
  for (int k = ty; k < 32; k += 8) M[(long)(i0 + k) * n + k0 + c0 + tx] = tile[tx][k];
}

// In-place inverse of the kBlk x kBlk block Pb[blockIdx.x] (column-major) by
// Gauss-Jordan elimination with partial (row) pivoting in LDS: at step j the
// largest |A(i, j)|, i >= j, is swapped into row j, row j is divided by the
// pivot (its diagonal entry taking the pivot's inverse), every other row i
// loses A(i, j) times row j (its column-j entry taking -A(i, j) / pivot), and
// the row swaps are undone as column swaps in reverse order at the end.
// info[blockIdx.x] is set when a pivot is zero.
__global__ void __launch_bounds__(1024) k_gj_inv(float2* __restrict__ Pb, int* __restrict__ info) {
  constexpr int L = kBlk, LP = kBlk + 1;  // LDS column stride padded against bank conflicts
  __shared__ float2 A[L * LP];
  __shared__ float2 f[L];
  __shared__ int perm[L];
  __shared__ int sp;
  float2* P = Pb + (long)blockIdx.x * L * L;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < L * L; e += blockDim.x) A[(e / L) * LP + e % L] = P[e];
  __syncthreads();
  bool bad = false;
  for (int j = 0; j < L; ++j) {
    if (wave == 0) {
      float best = -1.f;
      int bi = j;
      for (int i = j + lane; i < L; i += 64) {
        const float2 v = A[j * LP + i];
        const float m = v.x * v.x + v.y * v.y;
        if (m > best) {
          best = m;
          bi = i;
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        const float ob = __shfl_xor(best, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (lane == 0) {
        sp = bi;
        perm[j] = bi;
      }
    }
    __syncthreads();
    const int p = sp;
    if (p != j && tid < L) {  // swap rows j and p (column tid)
      const float2 a = A[tid * LP + j];
      A[tid * LP + j] = A[tid * LP + p];
      A[tid * LP + p] = a;
    }
    __syncthreads();
    const float2 piv = A[j * LP + j];
    const float d = piv.x * piv.x + piv.y * piv.y;
    bad |= !(d > 0.f);
    const float2 ip = d > 0.f ? make_float2(piv.x / d, -piv.y / d) : make_float2(0.f, 0.f);
    if (tid < L) f[tid] = A[j * LP + tid];  // column j
    __syncthreads();
    if (tid < L) {  // row j divided by the pivot
      const float2 a = tid == j ? make_float2(1.f, 0.f) : A[tid * LP + j];
      A[tid * LP + j] = cmul(a, ip);
    }
    __syncthreads();
    for (int e = tid; e < L * L; e += blockDim.x) {
      const int i = e % L, c = e / L;
      if (i == j) continue;
      const float2 a = c == j ? make_float2(0.f, 0.f) : A[c * LP + i];
      A[c * LP + i] = csub(a, cmul(f[i], A[c * LP + j]));
    }
    __syncthreads();
  }
  for (int j = L - 1; j >= 0; --j) {
    const int p = perm[j];
    if (p != j && tid < L) {  // swap columns j and p (row tid)
      const float2 a = A[j * LP + tid];
      A[j * LP + tid] = A[p * LP + tid];
      A[p * LP + tid] = a;
    }
    __syncthreads();
  }
  for (int e = tid; e < L * L; e += blockDim.x) P[e] = A[(e / L) * LP + e % L];
  if (bad && tid == 0) info[blockIdx.x] = 1;
}

// kscale of every member set to 0 before k_tile_red_max.
__global__ void k_tile_red_zero(const int* __restrict__ ids, int cnt, float* __restrict__ sc) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) sc[ids[i]] = 0.f;
}

// Largest |W| entry of member blockIdx.y (the swept matrix at Kw + koff, side
// nl, holds -W in its lower triangle): work-group blockIdx.x covers one chunk
// of the n_t x n_t entries and folds its maximum into the member's kscale
// with an atomic max on the bits (the values are >= 0, whose order the bits
// keep).
__global__ void __launch_bounds__(256)
k_tile_red_max(const float2* __restrict__ Kw, const int* __restrict__ ncol, const unsigned short* __restrict__ cols,
               const int* __restrict__ ids, const long* __restrict__ koff, const int* __restrict__ nl, int T,
               float* __restrict__ sc) {
  const int i = blockIdx.y;
  const int tb = ids[i];
  const int nk = ncol[tb];
  const long nn = (long)nk * nk;
  const long e0 = (long)blockIdx.x * kBuildChunk;
  if (e0 >= nn) return;
  const long e1 = e0 + kBuildChunk < nn ? e0 + kBuildChunk : nn;
  const unsigned short* ct = cols + (long)tb * T * T;
  const float2* A = Kw + koff[i];
  const int n = nl[i];
  float mx = 0.f;
  for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int p = (int)(e % nk), q = (int)(e / nk);
    if (p < q || ct[p] == kPadCol || ct[q] == kPadCol) continue;  // the lower triangle holds -W
    const float2 v = A[(long)q * n + p];
    mx = fmaxf(mx, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  __shared__ float red[4];
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float m = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    if (m > 0.f) atomicMax(reinterpret_cast<unsigned int*>(sc + tb), __float_as_uint(m));
  }
}

// W16 of member blockIdx.y: W / scale in FP16 (W = minus the swept matrix),
// the blocks (bi, bj <= bi) of its lower triangle (kBlkW x kBlkW each, column-major, zero beyond
// n_t) in row-major block order at woff[tb]; work-group blockIdx.x covers one
// chunk.  The matrix side nl is a multiple of kBlkW (= kBlk).
__global__ void __launch_bounds__(256)
k_tile_red_pack(const float2* __restrict__ Kw, const int* __restrict__ ncol, const unsigned short* __restrict__ cols,
                const int* __restrict__ ids, const long* __restrict__ koff, const int* __restrict__ nl,
                const long* __restrict__ woff, const float* __restrict__ sc, __half2* __restrict__ W16, int T) {
  constexpr int B = kBlkW;
  const int i = blockIdx.y;
  const int tb = ids[i];
  const int nk = ncol[tb];
  const int n = nl[i];
  const int nB = n / B;
  const long total = (long)nB * (nB + 1) / 2 * B * B;
  const long e0 = (long)blockIdx.x * kBuildChunk;
  if (e0 >= total) return;
  const long e1 = e0 + kBuildChunk < total ? e0 + kBuildChunk : total;
  const unsigned short* ct = cols + (long)tb * T * T;
  const float2* A = Kw + koff[i];
  const float r = sc[tb] > 0.f ? 1.f / sc[tb] : 0.f;  // 0: W_t = 0
  __half2* Wt = W16 + woff[tb];
  for (long e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int b = (int)(e / (B * B)), w = (int)(e % (B * B));
    int bi = (int)((sqrtf(8.f * b + 1.f) - 1.f) * 0.5f);  // block row: bi (bi + 1) / 2 <= b
    while (bi * (bi + 1) / 2 > b) --bi;
    while ((bi + 1) * (bi + 2) / 2 <= b) ++bi;
    const int bj = b - bi * (bi + 1) / 2;
    const int p = bi * B + w % B, q = bj * B + w / B;
    float2 v = make_float2(0.f, 0.f);
    if (p < nk && q < nk && ct[p] != kPadCol && ct[q] != kPadCol) {
      const float2 a = p >= q ? A[(long)q * n + p] : A[(long)p * n + q];  // the sweep leaves -W, lower
      v = make_float2(-a.x * r, -a.y * r);
    }
    Wt[e] = __floats2half2_rn(v.x, v.y);
  }
}

// ---------------------------------------------------------------------------
// reduced form: application
// ---------------------------------------------------------------------------

// Two-component shuffle across the wave.
__device__ inline float2 shfl_xor2(float2 v, int o) {
  return make_float2(__shfl_xor(v.x, o, 64), __shfl_xor(v.y, o, 64));
}

// One stage of the transposing butterfly: the 2H values of every lane become
// H, each the sum of the lane's and its partner's (lane ^ o) value of the
// same column, the upper half kept by the lanes with bit o set.
template <int H>
__device__ inline void bfly_stage(float2* c, int lane, int o) {
  const bool up = lane & o;
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const float2 send = up ? c[k] : c[k + H];
    const float2 keep = up ? c[k + H] : c[k];
    c[k] = cadd(keep, shfl_xor2(send, o));
  }
}

// W_t is complex symmetric (W_t = (D_S^{-1} - G_SS)^{-1}, G symmetric), so
// only its lower triangle is stored, in kBlkW x kBlkW blocks.  A work-group
// takes block e = (i, j) of its system s's map block list and writes two
// partial products of kBlkW values (scaled W16): B x_j (for block
// row i) and, for j < i, B^T x_i (for block row j), x_i, x_j the tile's
// input at the cells of S in block rows i and j (io.v scaled by
// 1 / *(io.nrm + s sNrm bytes) when io.nrm is given).  Lane l of every wave
// holds rows 2l, 2l + 1 of the block (one 8-byte load per column), the four
// waves take 32 columns each; the row sums meet in LDS, and the 32 column
// sums of each wave are reduced across its lanes by a transposing butterfly
// (lane l ends with column l / 2).  Systems with *(done + s sd bytes) set
// are skipped.  The grid is (8 ceil(blocks / 8), systems): work-group L in
// dispatch order (on XCD L % 8) takes system (L / 8) % systems and block
// 8 (L / (8 systems)) + L % 8, so the systems of one block run back to back
// on one XCD and, on a shared map, its L2 serves the block's later reads
// (16 systems at N = 512 fetch 1.5 times W from HBM, not 16 times).  One
// system keeps block e = L.
template <int T>
__global__ void __launch_bounds__(256)
k_tile_ws(const __half2* __restrict__ W16, const long* __restrict__ woff, const int* __restrict__ ncol,
          const unsigned short* __restrict__ cols, const int* __restrict__ btab, const int* __restrict__ bstart,
          const int* __restrict__ bcount, int maps1, FnoIO io, float2* __restrict__ Pw, long spw,
          const int* __restrict__ done, long sd, int N) {
  constexpr int M2 = T * T, B = kBlkW;
  __shared__ float2 xj[B], xi[B];
  __shared__ float4 red[3][64];
  const int nt = N / T;
  const int nsys = gridDim.y;
  const long L = (long)blockIdx.y * gridDim.x + blockIdx.x;
  const long q = L >> 3;
  const int s = (int)(q % nsys), map = maps1 ? 0 : s, e = (int)((q / nsys) * 8 + (L & 7));
  if (e >= bcount[map]) return;
  if (done && *boff(done, (long)s * sd)) return;
  const int pk = btab[bstart[map] + e];
  const int t = pk >> 16, i = (pk >> 8) & 255, j = pk & 255;
  const long tb = (long)map * nt * nt + t;
  const int n = ncol[tb];
  const unsigned short* ct = cols + tb * M2;
  float sc = 1.f;
  if (io.nrm) {
    const double d = *boff(io.nrm, (long)s * io.sNrm);
    sc = d > 0.0 ? (float)(1.0 / d) : 0.f;
  }
  {
    const int h = threadIdx.x >> 7, r = threadIdx.x & (B - 1);
    const int p = (h ? i : j) * B + r;
    float2 xv = make_float2(0.f, 0.f);
    if (p < n) {
      const int c = ct[p];
      if (c != kPadCol) {
        xv = io.v[(long)s * io.sV + tile_cell(t, c, nt, T, N)];
        xv.x *= sc;
        xv.y *= sc;
      }
    }
    (h ? xi : xj)[r] = xv;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = 2 * lane, q0 = 32 * wave;
  typedef unsigned int u2v_t __attribute__((ext_vector_type(2)));
  const u2v_t* Bp =
      reinterpret_cast<const u2v_t*>(W16 + woff[tb] + (long)(i * (i + 1) / 2 + j) * B * B + r0);
  const float2 x0 = xi[r0], x1 = xi[r0 + 1];
  const bool off = i != j;  // uniform over the work-group
  float2* P = Pw + (long)s * spw + (long)e * 2 * B;
  float ax = 0.f, ay = 0.f, bx = 0.f, by = 0.f;
#pragma unroll 1
  for (int g = 0; g < 2; ++g) {
    uint2 w[16];  // each vector load is copied out at once (rows r0 and r0 + 1)
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const u2v_t l = __builtin_nontemporal_load(Bp + (long)(q0 + 16 * g + u) * (B / 2));
      w[u] = make_uint2(l.x, l.y);
    }
    float2 c[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const float2 xv = xj[q0 + 16 * g + u];
      const float2 k0 = __half22float2(__builtin_bit_cast(__half2, w[u].x));
      const float2 k1 = __half22float2(__builtin_bit_cast(__half2, w[u].y));
      ax += k0.x * xv.x - k0.y * xv.y;
      ay += k0.x * xv.y + k0.y * xv.x;
      bx += k1.x * xv.x - k1.y * xv.y;
      by += k1.x * xv.y + k1.y * xv.x;
      c[u] = cadd(cmul(k0, x0), cmul(k1, x1));
    }
    if (off) {  // the 16 column sums of B^T x_i over the wave: lane l ends with column l / 4
      bfly_stage<8>(c, lane, 32);
      bfly_stage<4>(c, lane, 16);
      bfly_stage<2>(c, lane, 8);
      bfly_stage<1>(c, lane, 4);
      c[0] = cadd(c[0], shfl_xor2(c[0], 2));
      c[0] = cadd(c[0], shfl_xor2(c[0], 1));
      if ((lane & 3) == 0) P[B + q0 + 16 * g + (lane >> 2)] = c[0];
    }
  }
  if (wave > 0) red[wave - 1][lane] = make_float4(ax, ay, bx, by);
  __syncthreads();
  if (wave > 0) return;
#pragma unroll
  for (int w = 0; w < 3; ++w) {
    const float4 o = red[w][lane];
    ax += o.x;
    ay += o.y;
    bx += o.z;
    by += o.w;
  }
  P[r0] = make_float2(ax, ay);
  P[r0 + 1] = make_float2(bx, by);
}

// u_t of block row blockIdx.x % nbmax of tile blockIdx.x / nbmax, system
// blockIdx.y: the partial products of the blocks (i, j <= i) of the row and
// (k > i, i) of the column, summed in that fixed order, times the tile's
// scale, written to U at the cells of S.
template <int T>
__global__ void __launch_bounds__(kBlkW)
k_tile_wsum(const float* __restrict__ ksc, const int* __restrict__ ncol, const unsigned short* __restrict__ cols,
            const int* __restrict__ eloc, int maps1, int nbmax, const float2* __restrict__ Pw, long spw,
            float2* __restrict__ U, const int* __restrict__ done, long sd, int N) {
  constexpr int M2 = T * T, B = kBlkW;
  const int nt = N / T;
  const int s = blockIdx.y, map = maps1 ? 0 : s;
  const int t = blockIdx.x / nbmax, i = blockIdx.x % nbmax;
  const long tb = (long)map * nt * nt + t;
  const int n = ncol[tb], nB = (n + B - 1) / B;
  if (i >= nB) return;
  if (done && *boff(done, (long)s * sd)) return;
  const int r = threadIdx.x, p = i * B + r;
  const float2* P = Pw + (long)s * spw + (long)eloc[tb] * 2 * B;
  float2 acc = make_float2(0.f, 0.f);
  for (int j = 0; j <= i; ++j) acc = cadd(acc, P[(long)(i * (i + 1) / 2 + j) * 2 * B + r]);
  for (int k = i + 1; k < nB; ++k) acc = cadd(acc, P[(long)(k * (k + 1) / 2 + i) * 2 * B + B + r]);
  if (p < n) {
    const int c = cols[tb * M2 + p];
    if (c != kPadCol) {
      const float m = ksc[tb];
      U[((long)s * nt * nt + t) * M2 + c] = make_float2(acc.x * m, acc.y * m);
    }
  }
}

// z = x + Y (Y in tile-major order) for system blockIdx.y; x is io.v scaled
// as in k_tile_ws and is written to io.vw when io.nrm and io.vw are given.
template <int T>
__global__ void __launch_bounds__(256)
k_tile_combine(FnoIO io, const float2* __restrict__ Y, const int* __restrict__ done, long sd, int N) {
  const int s = blockIdx.y;
  if (done && *boff(done, (long)s * sd)) return;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)N * N) return;
  float sc = 1.f;
  if (io.nrm) {
    const double d = *boff(io.nrm, (long)s * io.sNrm);
    sc = d > 0.0 ? (float)(1.0 / d) : 0.f;
  }
  const int i = (int)(e / N), j = (int)(e % N);
  const int nt = N / T;
  float2 xv = io.v[(long)s * io.sV + e];
  xv.x *= sc;
  xv.y *= sc;
  if (io.nrm && io.vw) io.vw[(long)s * io.sV + e] = xv;
  const float2 g = Y[((long)s * nt * nt + (i / T) * nt + j / T) * (T * T) + (i % T) * T + j % T];
  io.z[(long)s * io.sZ + e] = cadd(xv, g);
}

int ensure_scratch(cgm_tiles* t, int nsys) {
  if (nsys <= t->scratch_sys) return CGM_SUCCESS;
  const size_t n = (size_t)t->N * t->N;
  const size_t tm = (size_t)t->nt * t->nt * t->T * t->P.M;
  void* ptrs[] = {t->U, t->Y, t->Tm};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  t->bytes -= (2 * n + tm) * sizeof(float2) * t->scratch_sys;
  t->U = t->Y = t->Tm = nullptr;
  t->scratch_sys = 0;
  if (hipMalloc((void**)&t->U, sizeof(float2) * n * nsys) != hipSuccess ||
      hipMalloc((void**)&t->Y, sizeof(float2) * n * nsys) != hipSuccess ||
      hipMalloc((void**)&t->Tm, sizeof(float2) * tm * nsys) != hipSuccess) {
    cgm_set_last_error("tile preconditioner: scratch allocation for %d systems failed", nsys);
    return CGM_ERROR_ALLOC;
  }
  // the cells with chi = 0 keep u = 0: the application writes only the cells of S
  if (hipMemset(t->U, 0, sizeof(float2) * n * nsys) != hipSuccess) {
    cgm_set_last_error("tile preconditioner: scratch clear failed");
    return CGM_ERROR_DEVICE;
  }
  t->scratch_sys = nsys;
  t->bytes += (2 * n + tm) * sizeof(float2) * nsys;
  return grow(t, &t->Pw, &t->pw_cap, (size_t)nsys * t->bcount_max * 2 * kBlkW);
}

// ev, when given: events recorded around k_tile_ws (live timing).
int launch_apply_red(const cgm_tiles* tl, hipStream_t st, const FnoIO& io, const int* done, long sd, int nsys,
                     hipEvent_t* ev) {
  cgm_tiles* t = const_cast<cgm_tiles*>(tl);  // the apply scratch grows with the batch
  int rc = ensure_scratch(t, nsys);
  if (rc) return rc;
  const int ntiles = tl->nt * tl->nt;
  const int B = nsys * ntiles;
  if (B > 65535) {
    cgm_set_last_error("%d systems exceed the tile convolution's batch limit", nsys);
    return CGM_ERROR_UNSUPPORTED;
  }
  const int maps1 = tl->maps == 1 ? 1 : 0;
  const long spw = (long)tl->bcount_max * 2 * kBlkW;
  if (tl->bcount_max > 0) {
    if (ev) CGM_HIP(hipEventRecord(ev[0], st));
    if (tl->T == 64)
      hipLaunchKernelGGL((k_tile_ws<64>), dim3((tl->bcount_max + 7) / 8 * 8, nsys), dim3(256), 0, st, tl->W16, tl->woff, tl->ncol,
                         tl->cols, tl->btab, tl->bstart, tl->bcount, maps1, io, tl->Pw, spw, done, sd, tl->N);
    else
      hipLaunchKernelGGL((k_tile_ws<128>), dim3((tl->bcount_max + 7) / 8 * 8, nsys), dim3(256), 0, st, tl->W16, tl->woff, tl->ncol,
                         tl->cols, tl->btab, tl->bstart, tl->bcount, maps1, io, tl->Pw, spw, done, sd, tl->N);
    CGM_HIP(hipGetLastError());
    if (ev) CGM_HIP(hipEventRecord(ev[1], st));
    if (tl->T == 64)
      hipLaunchKernelGGL((k_tile_wsum<64>), dim3(ntiles * tl->nbmax, nsys), dim3(kBlkW), 0, st, tl->kscale, tl->ncol,
                         tl->cols, tl->eloc, maps1, tl->nbmax, tl->Pw, spw, tl->U, done, sd, tl->N);
    else
      hipLaunchKernelGGL((k_tile_wsum<128>), dim3(ntiles * tl->nbmax, nsys), dim3(kBlkW), 0, st, tl->kscale, tl->ncol,
                         tl->cols, tl->eloc, maps1, tl->nbmax, tl->Pw, spw, tl->U, done, sd, tl->N);
    CGM_HIP(hipGetLastError());
  } else if (ev) {  // no cell of chi != 0: nothing streamed
    CGM_HIP(hipEventRecord(ev[0], st));
    CGM_HIP(hipEventRecord(ev[1], st));
  }
  BatchStrides bs;
  bs.x = tl->M2;
  bs.y = tl->M2;
  bs.T = (long)tl->T * tl->P.M;
  rc = launch_pass1<float>(tl->P, st, tl->U, nullptr, nullptr, nullptr, kModeGreen, tl->Tm, tl->tw32, nullptr, B, bs);
  if (!rc) rc = launch_pass2<float, float2>(tl->P, st, tl->Tm, tl->gq, tl->tw32, nullptr, B, bs);
  if (!rc)
    rc = launch_pass3<float>(tl->P, st, tl->Tm, tl->U, nullptr, nullptr, tl->Y, kModeGreen, tl->tw32, nullptr, B, bs);
  if (rc) {
    cgm_set_last_error("tile convolution at T = %d failed", tl->T);
    return rc;
  }
  const long n = (long)tl->N * tl->N;
  if (tl->T == 64)
    hipLaunchKernelGGL((k_tile_combine<64>), dim3((n + 255) / 256, nsys), dim3(256), 0, st, io, tl->Y, done, sd,
                       tl->N);
  else
    hipLaunchKernelGGL((k_tile_combine<128>), dim3((n + 255) / 256, nsys), dim3(256), 0, st, io, tl->Y, done, sd,
                       tl->N);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

// The storage of W of maps 0 .. maps-1 from n_t (nc, [maps][tiles]): block
// lists, offsets, entries streamed per map; uploads them and sizes W16 and the
// partial products.
int layout_reduced(cgm_tiles* t, const std::vector<int>& nc, int maps) {
  hipStream_t st = t->ctx->stream;
  const int ntiles = t->nt * t->nt;
  const int nb = ntiles * maps;
  constexpr long BB = (long)kBlkW * kBlkW;
  std::vector<long> off(nb);
  std::vector<int> eloc(nb), btab, bstart(maps), bcount(maps);
  long total = 0;
  t->kept.assign(maps, 0.0);
  t->nbmax = 0;
  t->bcount_max = 0;
  for (int m = 0; m < maps; ++m) {
    bstart[m] = (int)btab.size();
    int le = 0;
    for (int tl = 0; tl < ntiles; ++tl) {
      const int k = m * ntiles + tl;
      const int nB = (nc[k] + kBlkW - 1) / kBlkW;
      const int nblk = nB * (nB + 1) / 2;
      off[k] = total;
      total += nblk * BB;
      t->kept[m] += (double)nblk * BB;
      eloc[k] = le;
      le += nblk;
      t->nbmax = std::max(t->nbmax, nB);
      for (int i = 0; i < nB; ++i)
        for (int j = 0; j <= i; ++j) btab.push_back(tl << 16 | i << 8 | j);
    }
    bcount[m] = le;
    t->bcount_max = std::max(t->bcount_max, le);
  }
  int rc;
  if ((rc = grow(t, &t->W16, &t->w16_cap, (size_t)std::max(total, 1L)))) return rc;
  if ((rc = grow(t, &t->btab, &t->btab_cap, std::max<size_t>(btab.size(), 1)))) return rc;
  if ((rc = grow(t, &t->Pw, &t->pw_cap, (size_t)t->scratch_sys * t->bcount_max * 2 * kBlkW))) return rc;
  CGM_HIP(hipMemcpyAsync(t->woff, off.data(), sizeof(long) * nb, hipMemcpyHostToDevice, st));
  CGM_HIP(hipMemcpyAsync(t->eloc, eloc.data(), sizeof(int) * nb, hipMemcpyHostToDevice, st));
  if (!btab.empty())
    CGM_HIP(hipMemcpyAsync(t->btab, btab.data(), sizeof(int) * btab.size(), hipMemcpyHostToDevice, st));
  CGM_HIP(hipMemcpyAsync(t->bstart, bstart.data(), sizeof(int) * maps, hipMemcpyHostToDevice, st));
  CGM_HIP(hipMemcpyAsync(t->bcount, bcount.data(), sizeof(int) * maps, hipMemcpyHostToDevice, st));
  CGM_HIP(hipStreamSynchronize(st));  // the host vectors are read by the copies above
  return CGM_SUCCESS;
}

int upload_twiddles(cgm_tiles* t) {
  const int M = t->P.M;
  std::vector<float2> t32(M);
  std::vector<double2> t64(M);
  for (int q = 0; q < M; ++q) {
    const double a = -2.0 * CGM_PI * (double)q / (double)M;
    t64[q] = make_double2(cos(a), sin(a));
    t32[q] = make_float2((float)t64[q].x, (float)t64[q].y);
  }
  CGM_HIP(hipMemcpy(t->tw32, t32.data(), sizeof(float2) * M, hipMemcpyHostToDevice));
  CGM_HIP(hipMemcpy(t->tw64, t64.data(), sizeof(double2) * M, hipMemcpyHostToDevice));
  return CGM_SUCCESS;
}

// Symmetric block sweep (Gauss-Jordan on a symmetric matrix), in place, of
// the members' matrices A = D_S^{-1} - G_SS (side nl, column-major at
// Kw + koff, only the lower triangle kept): for every block column k0
// (kBlk wide) the pivot block is inverted (k_gj_inv, batched over the members
// still that large), then with P = A_kk^{-1} and C the block column k0 of the
// symmetric matrix,
//     A -= (C P) C^T (lower triangle, rocBLAS csyrkx);  A[:, k0 cols] = C P
//     (rows outside the pivot; the rows above land in the block row);  A_kk = -P.
// The matrix stays symmetric at every step (the sweep operator), and after
// the last block it holds -A^{-1} = -W.  No pivoting between blocks: the pivot
// blocks are Schur complements of the tile operator's restrictions to groups
// of cells, which the near-resonant but physical operators keep invertible (a
// zero pivot is reported per member).  In complex64 on the 128 x 128 tiles of
// the benchmark maps the result is within 3.5e-5 - 5.5e-5 of a complex128
// inverse (the unsymmetric block Gauss-Jordan: 3e-5 - 7e-5; LU with partial
// pivoting: 1e-5 - 2e-5), far below the FP16 storage of W, for half the
// flops of the unsymmetric form.
// Frees the FP32 matrices and blocks of the symmetric block sweep.
void release_factor_workspace(cgm_tiles* t) {
  float2** bufs[] = {&t->Kw, &t->Pb, &t->Rw, &t->Cw};
  size_t* caps[] = {&t->kw_cap, &t->pb_cap, &t->rw_cap, &t->cw_cap};
  for (int i = 0; i < 4; ++i) {
    if (*bufs[i]) (void)hipFree(*bufs[i]);
    t->bytes -= *caps[i] * sizeof(float2);
    *bufs[i] = nullptr;
    *caps[i] = 0;
  }
}

// *bad: the first member with a zero pivot, or -1.
int gj_invert(cgm_tiles* t, int members, const std::vector<int>& nlh, const std::vector<long>& koffh, int* bad) {
  hipStream_t st = t->ctx->stream;
  const int nmx = nlh[0];  // members are sorted by size, largest first
  int rc;
  if ((rc = grow(t, &t->Pb, &t->pb_cap, (size_t)members * kBlk * kBlk))) return rc;
  if ((rc = grow(t, &t->Rw, &t->rw_cap, (size_t)kBlk * nmx))) return rc;
  if ((rc = grow(t, &t->Cw, &t->cw_cap, (size_t)kBlk * nmx))) return rc;
  CGM_HIP(hipMemsetAsync(t->info, 0, sizeof(int) * members, st));
  const rocblas_float_complex one{1.f, 0.f}, mone{-1.f, 0.f}, zero{0.f, 0.f};
  const rocblas_operation nn = rocblas_operation_none;
  for (int k0 = 0; k0 < nmx; k0 += kBlk) {
    int cnt = 0;
    while (cnt < members && nlh[cnt] > k0) ++cnt;
    hipLaunchKernelGGL(k_gj_gather, dim3(cnt), dim3(256), 0, st, t->Kw, t->koff, t->nl, k0, t->Pb);
    hipLaunchKernelGGL(k_gj_inv, dim3(cnt), dim3(1024), 0, st, t->Pb, t->info);
    CGM_HIP(hipGetLastError());
    for (int i = 0; i < cnt; ++i) {
      const int n = nlh[i];
      float2* M = t->Kw + koffh[i];
      float2* P = t->Pb + (long)i * kBlk * kBlk;
      const dim3 tiles32(n / 32, kBlk / 32);
      hipLaunchKernelGGL(k_sw_gather, tiles32, dim3(256), 0, st, M, n, k0, t->Cw);
      CGM_HIP(hipGetLastError());
      rocblas_float_complex* Mr = reinterpret_cast<rocblas_float_complex*>(M);
      rocblas_float_complex* Pr = reinterpret_cast<rocblas_float_complex*>(P);
      rocblas_float_complex* C = reinterpret_cast<rocblas_float_complex*>(t->Cw);
      rocblas_float_complex* CP = reinterpret_cast<rocblas_float_complex*>(t->Rw);
      rocblas_status s = rocblas_cgemm(t->h, nn, nn, n, kBlk, kBlk, &one, C, n, Pr, kBlk, &zero, CP, n);
      if (s == rocblas_status_success)  // lower triangle of M -= (C P) C^T
        s = rocblas_csyrkx(t->h, rocblas_fill_lower, nn, n, kBlk, &mone, CP, n, C, n, &one, Mr, n);
      CGM_CHECK(s == rocblas_status_success, CGM_ERROR_DEVICE, "rocblas cgemm / csyrkx: status %d", (int)s);
      hipLaunchKernelGGL(k_sw_scatter, tiles32, dim3(256), 0, st, M, n, k0, t->Rw, P);
      CGM_HIP(hipGetLastError());
    }
  }
  std::vector<int> info(members);
  CGM_HIP(hipMemcpyAsync(info.data(), t->info, sizeof(int) * members, hipMemcpyDeviceToHost, st));
  CGM_HIP(hipStreamSynchronize(st));
  *bad = -1;
  for (int i = 0; i < members && *bad < 0; ++i)
    if (info[i]) *bad = i;
  return CGM_SUCCESS;
}

int factor_reduced(cgm_tiles* t, const double2* blk, const float* chi, long chi_stride, int maps) {
  cgm_context* c = t->ctx;
  hipStream_t st = c->stream;
  const int ntiles = t->nt * t->nt;
  const int nb = ntiles * maps;
  const int T = t->T, N = t->N;
  // spectrum of the tile convolution from the T x T corner of the block
  CGM_HIP(hipMemcpy2DAsync(t->blkT, sizeof(double2) * T, blk, sizeof(double2) * N, sizeof(double2) * T, T,
                           hipMemcpyDeviceToDevice, st));
  int rc = spectrum_from_block(t->P, st, t->blkT, t->tw64, t->gq, nullptr);
  if (rc) return rc;
  if (T == 64)
    hipLaunchKernelGGL((k_tile_cols<64>), dim3(ntiles, maps), dim3(1024), 0, st, chi, chi_stride, N, t->ncol,
                       t->cols);
  else
    hipLaunchKernelGGL((k_tile_cols<128>), dim3(ntiles, maps), dim3(1024), 0, st, chi, chi_stride, N, t->ncol,
                       t->cols);
  CGM_HIP(hipGetLastError());
  std::vector<int> nc(nb);
  CGM_HIP(hipMemcpyAsync(nc.data(), t->ncol, sizeof(int) * nb, hipMemcpyDeviceToHost, st));
  CGM_HIP(hipStreamSynchronize(st));
  // the storage of W and the factor members (tiles with n_t > 0, largest first)
  if ((rc = layout_reduced(t, nc, maps))) return rc;
  std::vector<int> mem;
  for (int k = 0; k < nb; ++k)
    if (nc[k] > 0) mem.push_back(k);
  std::stable_sort(mem.begin(), mem.end(), [&](int x, int y) { return nc[x] > nc[y]; });
  const int members = (int)mem.size();
  // u = 0 on the cells with chi = 0 of the new maps: the application writes only the cells of S
  CGM_HIP(hipMemsetAsync(t->U, 0, sizeof(float2) * (size_t)N * N * t->scratch_sys, st));
  if (members == 0) {
    CGM_HIP(hipStreamSynchronize(st));
    return CGM_SUCCESS;
  }
  std::vector<int> nlh(members);
  std::vector<long> koffh(members);
  long ktot = 0;
  int chunks = 1;
  for (int i = 0; i < members; ++i) {
    nlh[i] = (nc[mem[i]] + kBlk - 1) / kBlk * kBlk;
    koffh[i] = ktot;
    ktot += (long)nlh[i] * nlh[i];
    chunks = std::max(chunks, (int)(((long)nlh[i] * nlh[i] + kBuildChunk - 1) / kBuildChunk));
  }
  if ((rc = grow(t, &t->Kw, &t->kw_cap, (size_t)ktot))) return rc;
  CGM_HIP(hipMemcpyAsync(t->ids, mem.data(), sizeof(int) * members, hipMemcpyHostToDevice, st));
  CGM_HIP(hipMemcpyAsync(t->nl, nlh.data(), sizeof(int) * members, hipMemcpyHostToDevice, st));
  CGM_HIP(hipMemcpyAsync(t->koff, koffh.data(), sizeof(long) * members, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_tile_build_red, dim3(chunks, members), dim3(256), 0, st, blk, chi, chi_stride, t->ncol, t->cols,
                     t->ids, t->koff, t->nl, t->Kw, N, T);
  CGM_HIP(hipGetLastError());
  int bad = -1;
  if ((rc = gj_invert(t, members, nlh, koffh, &bad))) return rc;
  if (bad >= 0) {
    const int tb = mem[bad];
    cgm_set_last_error("tile %d of map %d: a pivot block of its reduced operator is singular", tb % ntiles,
                       tb / ntiles);
    return CGM_ERROR_NONFINITE;
  }
  hipLaunchKernelGGL(k_tile_red_zero, dim3(1), dim3(256), 0, st, t->ids, members, t->kscale);
  hipLaunchKernelGGL(k_tile_red_max, dim3(chunks, members), dim3(256), 0, st, t->Kw, t->ncol, t->cols, t->ids,
                     t->koff, t->nl, T, t->kscale);
  hipLaunchKernelGGL(k_tile_red_pack, dim3(chunks, members), dim3(256), 0, st, t->Kw, t->ncol, t->cols, t->ids,
                     t->koff, t->nl, t->woff, t->kscale, t->W16, T);
  CGM_HIP(hipGetLastError());
  CGM_HIP(hipStreamSynchronize(st));  // the host vectors are read by the copies above
  release_factor_workspace(t);  // FP32 matrices: twice the bytes of W, needed only here
  return CGM_SUCCESS;
}

}  // namespace

namespace cgm {

cgm_context* tiles_context(const cgm_tiles* t) { return t->ctx; }
int tiles_maps(const cgm_tiles* t) { return t->maps; }

// Algorithmic bytes of one application to nsys systems: the inverse entries
// streamed (dense form: every entry of every tile in FP32, once per system or
// once per group of four right-hand sides of one map; reduced form: the FP16
// entries of W of the map each system uses) plus v in and z out.
double apply_bytes(const cgm_tiles* tl, int nsys) {
  double e = 0.0;
  if (tl->reduced) {
    for (int s = 0; s < nsys; ++s) e += tl->kept[tl->maps == 1 ? 0 : s] * 4.0;
  } else {
    const int streams = (tl->maps == 1 && nsys > 1) ? (nsys + 3) / 4 : nsys;
    if (tl->maps == 1)
      e = tl->kept[0] * 8.0 * streams;
    else
      for (int s = 0; s < nsys; ++s) e += tl->kept[s] * 8.0;
  }
  return e + 16.0 * tl->N * tl->N * nsys;
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
  // streaming kernel (k_tile_apply, or k_tile_ws of the reduced form) and the
  // application's algorithmic bytes
  LrTiming* tm = tl->ctx->lrt.on ? &tl->ctx->lrt : nullptr;
  const int sl = tm ? tm->cur : -1;
  constexpr int KK = LrTiming::kKernels;
  hipEvent_t* ev = sl >= 0 ? &tm->ev[2 * KK * sl + 4] : nullptr;
  if (ev && !tl->reduced) CGM_HIP(hipEventRecord(ev[0], st));
  int rc;
  switch (tl->M2) {
    case 64: rc = launch_apply_m<64>(tl, st, io, done, sd, nsys); break;
    case 256: rc = launch_apply_m<256>(tl, st, io, done, sd, nsys); break;
    case 1024: rc = launch_apply_m<1024>(tl, st, io, done, sd, nsys); break;
    case 4096:
    case 16384: rc = launch_apply_red(tl, st, io, done, sd, nsys, ev); break;
    default:
      cgm_set_last_error("tile side %d unsupported", tl->T);
      return CGM_ERROR_UNSUPPORTED;
  }
  if (ev && !rc) {
    if (!tl->reduced) CGM_HIP(hipEventRecord(ev[1], st));
    tm->slot_bytes[2][sl] = apply_bytes(tl, nsys);
    tm->slot_has[2][sl] = true;
  }
  return rc;
}

}  // namespace cgm

extern "C" int cgm_tiles_create(cgm_ctx_t c, int T, int batch, cgm_tiles_t* out) {
  CGM_CHECK(c && out, CGM_ERROR_INVALID_VALUE, "null argument");
  *out = nullptr;
  CGM_CHECK(T == 8 || T == 16 || T == 32 || T == 64 || T == 128, CGM_ERROR_UNSUPPORTED,
            "tile side %d: 8, 16, 32, 64 or 128 supported", T);
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
  t->reduced = T >= 64;
  const size_t ntiles = (size_t)t->nt * t->nt;
  const size_t M2 = (size_t)t->M2;
  bool ok = rocblas_create_handle(&t->h) == rocblas_status_success;
  size_t want = 0;
  auto alloc = [&](void** p, size_t b) {
    if (ok && hipMalloc(p, b) != hipSuccess) ok = false;
    want += b;
  };
  if (!t->reduced) {
    alloc((void**)&t->K, sizeof(float2) * M2 * M2 * ntiles * batch);
    alloc((void**)&t->ipiv, sizeof(int) * M2 * ntiles * batch);
    alloc((void**)&t->info, sizeof(int) * ntiles * batch);
  } else {
    const size_t nb = ntiles * batch;
    alloc((void**)&t->ncol, sizeof(int) * nb);
    alloc((void**)&t->cols, sizeof(unsigned short) * M2 * nb);
    alloc((void**)&t->woff, sizeof(long) * nb);
    alloc((void**)&t->kscale, sizeof(float) * nb);
    alloc((void**)&t->ids, sizeof(int) * nb);
    alloc((void**)&t->koff, sizeof(long) * nb);
    alloc((void**)&t->eloc, sizeof(int) * nb);
    alloc((void**)&t->bstart, sizeof(int) * batch);
    alloc((void**)&t->bcount, sizeof(int) * batch);
    alloc((void**)&t->nl, sizeof(int) * nb);
    alloc((void**)&t->info, sizeof(int) * nb);
    if (ok && plan_init(t->P, T) != CGM_SUCCESS) ok = false;
    const size_t M = (size_t)t->P.M;
    alloc((void**)&t->tw32, sizeof(float2) * M);
    alloc((void**)&t->tw64, sizeof(double2) * M);
    alloc((void**)&t->blkT, sizeof(double2) * M2);
    alloc((void**)&t->gq, sizeof(float2) * (M / 2 + 1) * (M / 2 + 1));
    if (ok && upload_twiddles(t) != CGM_SUCCESS) ok = false;
  }
  t->bytes = want;
  if (ok && t->reduced && ensure_scratch(t, batch) != CGM_SUCCESS) ok = false;
  if (!ok) {
    cgm_tiles_destroy(t);
    cgm_set_last_error("cgm_tiles_create: allocation failed (%zu MiB)", want >> 20);
    return CGM_ERROR_ALLOC;
  }
  *out = t;
  return CGM_SUCCESS;
}

extern "C" int cgm_tiles_destroy(cgm_tiles_t t) {
  if (!t) return CGM_SUCCESS;
  (void)hipSetDevice(t->device);
  void* ptrs[] = {t->K,    t->ipiv, t->info, t->ncol, t->cols, t->woff, t->kscale,
                  t->ids,  t->eloc, t->btab, t->bstart, t->bcount, t->Pw, t->koff, t->nl,
                  t->W16,  t->Kw,   t->Pb,    t->Rw,   t->Cw,
                  t->U,    t->Y,    t->Tm,   t->tw32, t->tw64, t->blkT,  t->gq};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (t->h) (void)rocblas_destroy_handle(t->h);
  delete t;
  return CGM_SUCCESS;
}

extern "C" int cgm_tiles_memory(cgm_tiles_t t, size_t* bytes) {
  CGM_CHECK(t && bytes, CGM_ERROR_INVALID_VALUE, "null argument");
  *bytes = t->bytes;
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
  t->kept.assign(maps, 0.0);
  CGM_CHECK(rocblas_set_stream(t->h, c->stream) == rocblas_status_success, CGM_ERROR_DEVICE,
            "rocblas_set_stream failed");
  if (t->reduced) {
    const int rc = factor_reduced(t, (const double2*)block, chi, maps == 1 ? 0 : chi_stride, maps);
    if (rc) return rc;
    CGM_HIP(hipStreamSynchronize(c->stream));
    t->maps = maps;
    return CGM_SUCCESS;
  }
  const int chunks = (int)(((long)M2 * M2 + kBuildChunk - 1) / kBuildChunk);
  hipLaunchKernelGGL(k_tile_build, dim3(ntiles * chunks, maps), dim3(256), 0, c->stream, (const double2*)block, chi,
                     chi_stride, t->K, t->N, t->T, chunks);
  CGM_HIP(hipGetLastError());
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
  t->kept.assign(maps, (double)ntiles * M2 * M2);
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
// Synthetic dense inverses of one map.
__global__ void k_tile_fill(float2* __restrict__ K, long total) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const float v = 1e-3f * (float)(e % 7);
    K[e] = make_float2(v, -v);
  }
}

// Synthetic reduced form of one map in which every cell carries chi != 0
// (n_t = T^2 for every tile): W16 (total entries), unit scales, identity
// cell lists.
__global__ void k_tile_fill_red(__half2* __restrict__ W16, long total, float* __restrict__ sc, int* __restrict__ ncol,
                                unsigned short* __restrict__ cols, int M2, long ntiles) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const float v = 1e-3f * (float)(e % 7);
    W16[e] = __floats2half2_rn(v, -v);
    if (e < ntiles) {
      sc[e] = 1.f;
      ncol[e] = M2;
    }
    if (e < ntiles * M2) cols[e] = (unsigned short)(e % M2);
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
  t->maps = 0;
  if (t->reduced) {
    const int rc = layout_reduced(t, std::vector<int>(ntiles, t->M2), 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_tile_fill_red, dim3(4096), dim3(256), 0, c->stream, t->W16, (long)t->kept[0], t->kscale,
                       t->ncol, t->cols, t->M2, ntiles);
  } else {
    hipLaunchKernelGGL(k_tile_fill, dim3(4096), dim3(256), 0, c->stream, t->K, total);
    t->kept.assign(1, (double)total);
  }
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
