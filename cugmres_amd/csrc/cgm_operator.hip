// Circulant-embedded BTTB operator on gfx950: kernel-spectrum precompute
// (FP64) and the three fused FFT passes of the matvec.
//
// Reference counterparts: Ghat = FFT(G_prep) at B/CUDA_GMRES/launch_GMRES.cuh:138-142;
// G_x_fft_matvec at B/CUDA_GMRES/Fast_matvec.cuh:1-56, i.e. the chain
// extend_by_zeros_kernel -> cufftExecC2C -> MatMul_ElemWise_kernel ->
// cufftExecC2C -> _2D_to_1D_kernel (B/CUDA_GMRES/GMRES_kernels.cuh:98-167,286-326).
//
// Here the embedding is M x M with M >= 2N-1 chosen FFT-friendly (2^a or
// 3*2^a), and the five reference launches collapse into three passes:
//   pass 1  row i < N:  u = chi .* (s x)[i, :] zero-padded to M, forward FFT
//                       along the row -> T[i, :]      (optionally writes s x back)
//   pass 2  column c:   forward FFT of T[:, c] (rows >= N are zero), multiply
//                       by Ghat[:, c], inverse FFT, keep rows < N -> T[:, c]
//   pass 3  row i < N:  inverse FFT of T[i, :], crop to N, scale 1/M^2 and
//                       combine with x (and b for a residual) -> y[i, :]
// T is N x M row-major (complex), so only N of the M rows are ever stored.
#include <math.h>
#include <string.h>

#include "cgm_fft.h"
#include "cgm_internal.h"
#include "cgm_special.h"

namespace cgm {

// ---------------------------------------------------------------------------
// pass kernels
// ---------------------------------------------------------------------------
// Work-group bound: float kernels accept tiles up to 1024 threads (autotuned
// rows/columns per work-group); double kernels keep the tighter bound that
// avoids register spills of the FP64 transforms.
template <typename T, int M> struct PassWG {
  static constexpr int value =
      (sizeof(T) == 4 && FftShape<M>::pow2) ? (FftShape<M>::WG > 1024 ? FftShape<M>::WG : 1024) : FftShape<M>::WG;
};

template <typename T, int M>
__global__ void __launch_bounds__((PassWG<T, M>::value))
k_pass1(const cplx<T>* __restrict__ x, cplx<T>* xw, const T* __restrict__ chi,
        const double* __restrict__ norm_ptr, int mode, cplx<T>* __restrict__ Tm, int N,
        int rows_per_wg, const cplx<T>* __restrict__ tw, const int* __restrict__ done, BatchStrides bs) {
  {
    const long k = blockIdx.y;
    x += k * bs.x;
    if (xw) xw += k * bs.x;
    if (chi) chi += k * bs.chi;
    Tm += k * bs.T;
    norm_ptr = boff(norm_ptr, k * bs.st);
    done = boff(done, k * bs.st);
  }
  const int dn = done ? *done : 0;  // tested once the row loads are in flight
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* smem = reinterpret_cast<cplx<T>*>(smem_raw);
  constexpr int NT = FftShape<M>::NT;
  constexpr int EPT = FftShape<M>::EPT;
  const int t = threadIdx.x % NT;
  const int rl = threadIdx.x / NT;
  const int row = blockIdx.x * rows_per_wg + rl;
  constexpr int ML = LdsLen<T, M>::value;
  cplx<T>* buf = smem + rl * ML;
  const bool live = row < N;
  const long rb = (long)row * N;
  const bool mul = mode == kModeForward || mode == kModeResidual;
  FftTwiddles<T, M> tws;
  tws.load(tw, t);
  cplx<T> xv[EPT];
  T cv[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = t + k * NT;
    xv[k] = (live && e < N) ? x[rb + e] : mk(T(0), T(0));
    cv[k] = (live && e < N && mul) ? chi[rb + e] : T(1);
  }
  if (dn) return;
  T s = T(1);
  if (norm_ptr) {
    const double nv = *norm_ptr;
    s = nv > 0.0 ? T(1.0 / nv) : T(0);
  }
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = t + k * NT;
    cplx<T> v = xv[k];
    if (norm_ptr) {
      v = cscale(v, s);
      if (xw && live && e < N) xw[rb + e] = v;
    }
    if (e < fft_lo_read_end<M>()) buf[lpad<T>(e)] = mul ? cscale(v, cv[k]) : v;
  }
  __syncthreads();
  lds_fft<T, M, -1, true>(buf, tws, t);  // the row is zero from N <= M/2 on
  if (live) {
    for (int e = t; e < M; e += NT) Tm[tpos(row, e, N)] = buf[lpad<T>(e)];
  }
}

template <typename T, typename G, int M>
__global__ void __launch_bounds__((PassWG<T, M>::value))
k_pass2(cplx<T>* __restrict__ Tm, const G* __restrict__ gq, int N, int cols_per_wg,
        const cplx<T>* __restrict__ tw, const int* __restrict__ done, BatchStrides bs) {
  Tm += (long)blockIdx.y * bs.T;
  done = boff(done, (long)blockIdx.y * bs.st);
  const int dn = done ? *done : 0;  // tested once the column loads are in flight
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* smem = reinterpret_cast<cplx<T>*>(smem_raw);
  constexpr int NT = FftShape<M>::NT;
  constexpr int H = M / 2 + 1;
  constexpr int ML = LdsLen<T, M>::value;
  const int grp = xcd_tile(blockIdx.x, gridDim.x);
  const int c0 = grp * cols_per_wg;
  const int nthr = blockDim.x;  // cols_per_wg * NT
  FftTwiddles<T, M> tws;
  tws.load(tw, threadIdx.x % NT);
  // gather the column tile: consecutive threads read consecutive columns.
  // Single precision: N <= M/2 + 1, so each thread gathers at most EPT/2 + 1
  // values into registers before the done test; double precision (whose
  // registers are spoken for) tests first and gathers straight to LDS.
  if constexpr (sizeof(T) == 4) {
    constexpr int GR = FftShape<M>::EPT / 2 + 1;
    cplx<T> gv[GR];
#pragma unroll
    for (int k = 0; k < GR; ++k) {
      const int idx = threadIdx.x + k * nthr;
      if (idx < N * cols_per_wg) {
        const int r = idx / cols_per_wg, c = idx - r * cols_per_wg;
        gv[k] = Tm[tpos(r, c0 + c, N)];
      }
    }
    if (dn) return;
#pragma unroll
    for (int k = 0; k < GR; ++k) {
      const int idx = threadIdx.x + k * nthr;
      if (idx < N * cols_per_wg) {
        const int r = idx / cols_per_wg, c = idx - r * cols_per_wg;
        smem[c * ML + lpad<T>(r)] = gv[k];
      }
    }
  } else {
    if (dn) return;
    for (int idx = threadIdx.x; idx < N * cols_per_wg; idx += nthr) {
      const int r = idx / cols_per_wg, c = idx - r * cols_per_wg;
      smem[c * ML + lpad<T>(r)] = Tm[tpos(r, c0 + c, N)];
    }
  }
  for (int idx = threadIdx.x; idx < (fft_lo_read_end<M>() - N) * cols_per_wg; idx += nthr) {
    const int r = N + idx / cols_per_wg, c = idx % cols_per_wg;
    smem[c * ML + lpad<T>(r)] = mk(T(0), T(0));
  }
  __syncthreads();
  const int t = threadIdx.x % NT;
  const int cl = threadIdx.x / NT;
  cplx<T>* buf = smem + cl * ML;
  lds_fft<T, M, -1, true>(buf, tws, t);  // rows N.. are zero
  const int col = c0 + cl;
  const int cq = col <= M - col ? col : M - col;
  const G* grow = gq + (long)cq * H;
  for (int e = t; e < M; e += NT) {
    const int fr = e <= M - e ? e : M - e;
    const G g = grow[fr];
    buf[lpad<T>(e)] = cmul(buf[lpad<T>(e)], mk(T(g.x), T(g.y)));
  }
  __syncthreads();
  lds_fft<T, M, +1, false, true>(buf, tws, t);  // rows < N are kept
  for (int idx = threadIdx.x; idx < N * cols_per_wg; idx += nthr) {
    const int r = idx / cols_per_wg, c = idx - r * cols_per_wg;
    Tm[tpos(r, c0 + c, N)] = smem[c * ML + lpad<T>(r)];
  }
}

// ---------------------------------------------------------------------------
// Pass 2 for M = 1024 in registers (complex64; chosen by the autotuner for
// batches).  1024 = 32 x 32: 32 lanes transform one column, each lane holding
// 32 values; a wave takes two columns and a work-group of four waves one
// panel of kPW = 8 columns.  Forward transform, element n = n1 + 32 n2 in
// lane n1, register n2:
//   A: Y[n1][k1] = DFT32 over n2 (registers), times W1024^(n1 k1);
//   exchange through LDS (lane k1 gathers Y[.][k1]);
//   B: X[k1 + 32 k2] = DFT32 over n1 (registers) in lane k1, register k2.
// The spectrum multiply runs on that layout, and the inverse retraces the
// steps with conjugate twiddles, ending with x[n1 + 32 n2] in lane n1.  One
// LDS exchange per direction (8 KB per column, XOR-swizzled so that both its
// writes and its reads are free of bank conflicts) replaces the eight
// LDS-staged radix stages of k_pass2.
// ---------------------------------------------------------------------------
__device__ __constant__ float2 kW32[22] = {
    {1.000000000f, 0.000000000f},   {0.980785280f, -0.195090322f},  {0.923879533f, -0.382683432f},
    {0.831469612f, -0.555570233f},  {0.707106781f, -0.707106781f},  {0.555570233f, -0.831469612f},
    {0.382683432f, -0.923879533f},  {0.195090322f, -0.980785280f},  {0.000000000f, -1.000000000f},
    {-0.195090322f, -0.980785280f}, {-0.382683432f, -0.923879533f}, {-0.555570233f, -0.831469612f},
    {-0.707106781f, -0.707106781f}, {-0.831469612f, -0.555570233f}, {-0.923879533f, -0.382683432f},
    {-0.980785280f, -0.195090322f}, {-1.000000000f, 0.000000000f},  {-0.980785280f, 0.195090322f},
    {-0.923879533f, 0.382683432f},  {-0.831469612f, 0.555570233f},  {-0.707106781f, 0.707106781f},
    {-0.555570233f, 0.831469612f}};  // exp(-2 pi i q / 32), q = 0..21

// v[k] <- sum_n v[n] exp(SIGN 2 pi i n k / 32): with n = 4 m + p, four 8-point
// DFTs over m, the twiddles W32^(p k1), then eight 4-point DFTs over p give
// X[k1 + 8 k2].
template <int SIGN>
__device__ inline void dft32(float2* v) {
  float2 a[4][8];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
#pragma unroll
    for (int m = 0; m < 8; ++m) a[p][m] = v[4 * m + p];
    dft8<float, SIGN>(a[p]);
  }
#pragma unroll
  for (int p = 1; p < 4; ++p)
#pragma unroll
    for (int k1 = 1; k1 < 8; ++k1) {
      float2 w = kW32[p * k1];
      if (SIGN > 0) w.y = -w.y;
      a[p][k1] = cmul(a[p][k1], w);
    }
#pragma unroll
  for (int k1 = 0; k1 < 8; ++k1) {
    float2 b[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    dft4<float, SIGN>(b);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 8 * k2] = b[k2];
  }
}

// Lane `lane` of a 32-lane column group: registers v[r] -> X[lane][r] in the
// group's LDS region L, then reads back v[r] = X[r][lane] (the transpose).
__device__ inline void xchg32(float2* __restrict__ L, float2* v, int lane) {
#pragma unroll
  for (int r = 0; r < 32; ++r) L[r * 32 + (lane ^ r)] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 32; ++r) v[r] = L[lane * 32 + (r ^ lane)];
  __syncthreads();
}

template <typename G>
__global__ void __launch_bounds__(256)
k_pass2w(float2* __restrict__ Tm, const G* __restrict__ gq, int N, const float2* __restrict__ tw,
         const int* __restrict__ done, BatchStrides bs) {
  constexpr int M = 1024;
  constexpr int H = M / 2 + 1;
  __shared__ float2 lds[8][1024];
  Tm += (long)blockIdx.y * bs.T;
  done = boff(done, (long)blockIdx.y * bs.st);
  const int panel = xcd_tile(blockIdx.x, gridDim.x);
  const int grp = threadIdx.x >> 5;  // column of the panel
  const int lane = threadIdx.x & 31;
  const int col = panel * kPW + grp;
  float2* L = lds[grp];
  // The panel (N rows x 8 columns, contiguous in tpos() layout) is staged in
  // LDS with a row stride of 9 values (coalesced global loads; the column
  // reads that follow fall in distinct banks), then each lane takes rows
  // n = lane + 32 r of its column (zero from N on).
  float2* S = &lds[0][0];
  constexpr int kRS = kPW + 1;
  const float2* src = Tm + ((long)panel * N) * kPW;
  for (int e = threadIdx.x; e < N * kPW; e += 256) S[(e / kPW) * kRS + e % kPW] = src[e];
  __syncthreads();
  if (done && *done) return;  // uniform over the work-group
  float2 v[32];
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    const int n = lane + 32 * r;
    v[r] = n < N ? S[n * kRS + grp] : make_float2(0.f, 0.f);
  }
  __syncthreads();  // the staging area is reused by the exchanges
  // forward: A over registers, twiddle, exchange, B
  dft32<-1>(v);
#pragma unroll
  for (int r = 1; r < 32; ++r) v[r] = cmul(v[r], tw[lane * r]);
  xchg32(L, v, lane);
  dft32<-1>(v);
  // lane k1 holds X[k1 + 32 k2] in v[k2]: multiply by the spectrum column
  const int cq = col <= M - col ? col : M - col;
  const G* grow = gq + (long)cq * H;
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    const int k = lane + 32 * r;
    const int fr = k <= M - k ? k : M - k;
    const G g = grow[fr];
    v[r] = cmul(v[r], make_float2((float)g.x, (float)g.y));
  }
  // inverse: A' over registers (k2), conjugate twiddle, exchange, B' (k1)
  dft32<1>(v);
#pragma unroll
  for (int r = 1; r < 32; ++r) {
    float2 w = tw[lane * r];
    w.y = -w.y;
    v[r] = cmul(v[r], w);
  }
  xchg32(L, v, lane);
  dft32<1>(v);
  // back through the same staging layout for coalesced stores
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    const int n = lane + 32 * r;
    if (n < N) S[n * kRS + grp] = v[r];
  }
  __syncthreads();
  float2* dst = Tm + ((long)panel * N) * kPW;
  for (int e = threadIdx.x; e < N * kPW; e += 256) dst[e] = S[(e / kPW) * kRS + e % kPW];
}

// Passes 1 and 3 for M = 1024 in registers, with the same 32 x 32 scheme:
// 32 lanes per row, eight rows per work-group.  Pass 1 loads row i of
// chi .* x (lane n1, register n2: element n1 + 32 n2), transforms it forward
// and writes T[i][k1 + 32 k2] from lane k1, register k2.  Pass 3 reads that
// layout, transforms it back, and lane n1 combines element n1 + 32 n2 of the
// row with x.  Only the plain operator modes (y = x - G(chi x), x - chi G(x),
// G x) without input scaling take these kernels.
__global__ void __launch_bounds__(256)
k_pass1w(const float2* __restrict__ x, const float* __restrict__ chi, int mul, float2* __restrict__ Tm, int N,
         const float2* __restrict__ tw, const int* __restrict__ done, BatchStrides bs) {
  __shared__ float2 lds[8][1024];
  {
    const long kb = blockIdx.y;
    x += kb * bs.x;
    if (chi) chi += kb * bs.chi;
    Tm += kb * bs.T;
    done = boff(done, kb * bs.st);
  }
  const int grp = threadIdx.x >> 5, lane = threadIdx.x & 31;
  const int row = blockIdx.x * 8 + grp;
  const bool live = row < N;
  const long rb = (long)row * N;
  float2 v[32];
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    const int n = lane + 32 * r;
    float2 e = make_float2(0.f, 0.f);
    if (live && n < N) {
      e = x[rb + n];
      if (mul) {
        const float c = chi[rb + n];
        e.x *= c;
        e.y *= c;
      }
    }
    v[r] = e;
  }
  if (done && *done) return;  // uniform over the work-group
  dft32<-1>(v);
#pragma unroll
  for (int r = 1; r < 32; ++r) v[r] = cmul(v[r], tw[lane * r]);
  xchg32(lds[grp], v, lane);
  dft32<-1>(v);
  if (live) {
#pragma unroll
    for (int r = 0; r < 32; ++r) Tm[tpos(row, lane + 32 * r, N)] = v[r];
  }
}

__global__ void __launch_bounds__(256)
k_pass3w(const float2* __restrict__ Tm, const float2* __restrict__ x, const float* __restrict__ chi,
         float2* __restrict__ y, int mode, int N, float inv, const float2* __restrict__ tw,
         const int* __restrict__ done, BatchStrides bs) {
  __shared__ float2 lds[8][1024];
  {
    const long kb = blockIdx.y;
    Tm += kb * bs.T;
    x += kb * bs.x;
    if (chi) chi += kb * bs.chi;
    y += kb * bs.y;
    done = boff(done, kb * bs.st);
  }
  const int grp = threadIdx.x >> 5, lane = threadIdx.x & 31;
  const int row = blockIdx.x * 8 + grp;
  const bool live = row < N;
  float2 v[32];
#pragma unroll
  for (int r = 0; r < 32; ++r) v[r] = live ? Tm[tpos(row, lane + 32 * r, N)] : make_float2(0.f, 0.f);
  if (done && *done) return;
  dft32<1>(v);
#pragma unroll
  for (int r = 1; r < 32; ++r) {
    float2 w = tw[lane * r];
    w.y = -w.y;
    v[r] = cmul(v[r], w);
  }
  xchg32(lds[grp], v, lane);
  dft32<1>(v);
  if (!live) return;
  const long rb = (long)row * N;
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    const int n = lane + 32 * r;
    if (n < N) {
      const float2 g = cscale(v[r], inv);
      float2 o;
      if (mode == kModeForward) o = csub(x[rb + n], g);
      else if (mode == kModeAdjoint) o = csub(x[rb + n], cscale(g, chi[rb + n]));
      else o = g;
      y[rb + n] = o;
    }
  }
}

template <typename T, int M>
__global__ void __launch_bounds__((PassWG<T, M>::value))
k_pass3(const cplx<T>* __restrict__ Tm, const cplx<T>* __restrict__ x, const T* __restrict__ chi,
        const cplx<T>* __restrict__ b, cplx<T>* __restrict__ y, int mode, int N, int rows_per_wg,
        T inv, const cplx<T>* __restrict__ tw, const int* __restrict__ done, BatchStrides bs) {
  {
    const long k = blockIdx.y;
    Tm += k * bs.T;
    x += k * bs.x;
    if (chi) chi += k * bs.chi;
    if (b) b += k * bs.b;
    y += k * bs.y;
    done = boff(done, k * bs.st);
  }
  if (done && *done) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* smem = reinterpret_cast<cplx<T>*>(smem_raw);
  constexpr int NT = FftShape<M>::NT;
  const int t = threadIdx.x % NT;
  const int rl = threadIdx.x / NT;
  const int row = blockIdx.x * rows_per_wg + rl;
  constexpr int ML = LdsLen<T, M>::value;
  cplx<T>* buf = smem + rl * ML;
  const bool live = row < N;
  FftTwiddles<T, M> tws;
  tws.load(tw, t);
  for (int e = t; e < M; e += NT) buf[lpad<T>(e)] = live ? Tm[tpos(row, e, N)] : mk(T(0), T(0));
  __syncthreads();
  lds_fft<T, M, +1, false, true>(buf, tws, t);  // columns < N are used
  if (!live) return;
  const long rb = (long)row * N;
  for (int e = t; e < N; e += NT) {
    const cplx<T> g = cscale(buf[lpad<T>(e)], inv);
    cplx<T> o;
    switch (mode) {
      case kModeForward: o = csub(x[rb + e], g); break;
      case kModeAdjoint: o = csub(x[rb + e], cscale(g, chi[rb + e])); break;
      case kModeGreen: o = g; break;
      case kModeResidual: o = csub(b[rb + e], csub(x[rb + e], g)); break;
      default: o = csub(b[rb + e], csub(x[rb + e], cscale(g, chi[rb + e]))); break;
    }
    y[rb + e] = o;
  }
}

// Full-size 2-D FFT helpers for the one-off spectrum precompute (FP64).
template <typename T, int M, int SIGN>
__global__ void __launch_bounds__(FftShape<M>::WG)
k_fft_rows(cplx<T>* __restrict__ data, int rows, int rows_per_wg, const cplx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* smem = reinterpret_cast<cplx<T>*>(smem_raw);
  constexpr int NT = FftShape<M>::NT;
  const int t = threadIdx.x % NT, rl = threadIdx.x / NT;
  const int row = blockIdx.x * rows_per_wg + rl;
  cplx<T>* buf = smem + rl * M;
  const bool live = row < rows;
  FftTwiddles<T, M> tws;
  tws.load(tw, t);
  for (int e = t; e < M; e += NT) buf[e] = live ? data[(long)row * M + e] : mk(T(0), T(0));
  __syncthreads();
  lds_fft<T, M, SIGN>(buf, tws, t);
  if (live)
    for (int e = t; e < M; e += NT) data[(long)row * M + e] = buf[e];
}

template <typename T, int M, int SIGN>
__global__ void __launch_bounds__(FftShape<M>::WG)
k_fft_cols(cplx<T>* __restrict__ data, int cols_per_wg, const cplx<T>* __restrict__ tw) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* smem = reinterpret_cast<cplx<T>*>(smem_raw);
  constexpr int NT = FftShape<M>::NT;
  const int c0 = blockIdx.x * cols_per_wg;
  for (int idx = threadIdx.x; idx < M * cols_per_wg; idx += blockDim.x) {
    const int r = idx / cols_per_wg, c = idx % cols_per_wg;
    smem[c * M + r] = data[(long)r * M + c0 + c];
  }
  __syncthreads();
  const int t = threadIdx.x % NT, cl = threadIdx.x / NT;
  FftTwiddles<T, M> tws;
  tws.load(tw, t);
  lds_fft<T, M, SIGN>(smem + cl * M, tws, t);
  for (int idx = threadIdx.x; idx < M * cols_per_wg; idx += blockDim.x) {
    const int r = idx / cols_per_wg, c = idx % cols_per_wg;
    data[(long)r * M + c0 + c] = smem[c * M + r];
  }
}

// ---------------------------------------------------------------------------
// Green's-function generators (FP64)
// ---------------------------------------------------------------------------
// Block g[i, j] (i, j < N) of the BTTB generator.
//  kind 0 (the benchmark operator, k0^2 G): off-diagonal kh^2 (i/4) H0(kh rho),
//         self term (i pi ka/2) H1(ka) - 1, ka = kh / sqrt(pi).
//  kind 1 (reference operator, B/MKL_GMRES_test_09122019/source/loop_functions.h:139-142):
//         (i/4) H0(k sqrt((i - sr)^2 + (j - sc)^2)) for every (i, j), the
//         half-cell shift sr = 0.5 (C++ orientation) or sc = 0.5 (Python
//         orientation, B/PyGMRES/GMRES.py:23-27).
__global__ void k_green_block(double2* __restrict__ blk, int N, int kind, double k, double sr,
                              double sc) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)N * N) return;
  const int i = (int)(p / N), j = (int)(p % N);
  double re, im, J0, J1, Y0, Y1;
  if (kind == 0) {
    if (i == 0 && j == 0) {
      const double ka = k / sqrt(CGM_PI);
      bessel01(ka, J0, J1, Y0, Y1);
      re = -0.5 * CGM_PI * ka * Y1 - 1.0;
      im = 0.5 * CGM_PI * ka * J1;
    } else {
      const double rho = sqrt((double)i * i + (double)j * j);
      bessel01(k * rho, J0, J1, Y0, Y1);
      re = -0.25 * k * k * Y0;
      im = 0.25 * k * k * J0;
    }
  } else {
    const double di = i - sr, dj = j - sc;
    bessel01(k * sqrt(di * di + dj * dj), J0, J1, Y0, Y1);
    re = -0.25 * Y0;
    im = 0.25 * J0;
  }
  blk[p] = make_double2(re, im);
}

// Reads the generator out of a (2N-1)^2 embedded array laid out as the
// reference's G_prep (B/MKL_GMRES_test_09122019/source/loop_functions.h:150-163).
__global__ void k_gprep_block(const float2* __restrict__ gprep, double2* __restrict__ blk, int N) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)N * N) return;
  const int i = (int)(p / N), j = (int)(p % N);
  const float2 v = gprep[(long)i * (2 * N - 1) + j];
  blk[p] = make_double2(v.x, v.y);
}

// E[(di mod M), (dj mod M)] = blk[|di|, |dj|] for |di|, |dj| < N, else 0.
__global__ void k_embed(const double2* __restrict__ blk, double2* __restrict__ E, int N, int M) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)M * M) return;
  const int r = (int)(p / M), c = (int)(p % M);
  const int ai = r < N ? r : (r > M - N ? M - r : -1);
  const int aj = c < N ? c : (c > M - N ? M - c : -1);
  E[p] = (ai >= 0 && aj >= 0) ? blk[(long)ai * N + aj] : make_double2(0.0, 0.0);
}

// Quarter of the even-symmetric spectrum, stored column-frequency major:
// gq[c' * (M/2+1) + r'] = Ghat[r', c'].
template <typename G>
__global__ void k_quarter(const double2* __restrict__ E, G* __restrict__ gq, int M) {
  const int H = M / 2 + 1;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long)H * H) return;
  const int cq = (int)(p / H), rq = (int)(p % H);
  const double2 v = E[(long)rq * M + cq];
  G o;
  o.x = v.x;
  o.y = v.y;
  gq[p] = o;
}

// ---------------------------------------------------------------------------
// host-side dispatch over the supported embedding sizes
// ---------------------------------------------------------------------------
#define CGM_FOR_EACH_M(X) \
  X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) \
  X(1024) X(1536) X(2048) X(3072) X(4096)

int fft_threads(int M) {
  switch (M) {
#define X(m) case m: return FftShape<m>::NT;
    CGM_FOR_EACH_M(X)
#undef X
  }
  return 0;
}

// Largest work-group (threads) the pass kernels of size M accept.
int pass_max_threads(int M, bool fp64) {
  switch (M) {
#define X(m) case m: return fp64 ? PassWG<double, m>::value : PassWG<float, m>::value;
    CGM_FOR_EACH_M(X)
#undef X
  }
  return 0;
}

int choose_embedding(int N) {
  static const int sizes[] = {16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768,
                              1024, 1536, 2048, 3072, 4096};
  for (int s : sizes)
    if (s >= 2 * N - 1) return s;
  return 0;
}

static int groups_for(int M, int target_threads) {
  const int nt = fft_threads(M);
  int g = target_threads / nt;
  return g < 1 ? 1 : g;
}

template <typename T>
int launch_pass1(const Plan& P, hipStream_t st, const cplx<T>* x, cplx<T>* xw, const T* chi,
                 const double* norm_ptr, int mode, cplx<T>* Tm, const cplx<T>* tw, const int* done,
                 int B, const BatchStrides& bs) {
  const int M = P.M, N = P.N;
  if constexpr (sizeof(T) == 4) {
    if (M == 1024 && P.wide1_batch > 0 && B >= P.wide1_batch && !xw && !norm_ptr &&
        (mode == kModeForward || mode == kModeAdjoint || mode == kModeGreen)) {
      hipLaunchKernelGGL(k_pass1w, dim3((N + 7) / 8, B), dim3(256), 0, st, (const float2*)x, (const float*)chi,
                         (int)(mode == kModeForward), (float2*)Tm, N, (const float2*)tw, done, bs);
      return hipGetLastError() == hipSuccess ? CGM_SUCCESS : CGM_ERROR_DEVICE;
    }
  }
  const int rows = P.rows1;
  const int grid = (N + rows - 1) / rows;
  const int threads = rows * fft_threads(M);
  const size_t lds = (size_t)rows * lds_len(M, sizeof(T) == 8) * sizeof(cplx<T>);
  if (threads > pass_max_threads(M, sizeof(T) == 8) || lds > 65536) return CGM_ERROR_UNSUPPORTED;
  switch (M) {
#define X(m)                                                                                 \
  case m:                                                                                    \
    hipLaunchKernelGGL((k_pass1<T, m>), dim3(grid, B), dim3(threads), lds, st, x, xw, chi,   \
                       norm_ptr, mode, Tm, N, rows, tw, done, bs);                            \
    break;
    CGM_FOR_EACH_M(X)
#undef X
    default: return CGM_ERROR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? CGM_SUCCESS : CGM_ERROR_DEVICE;
}

template <typename T, typename G>
int launch_pass2(const Plan& P, hipStream_t st, cplx<T>* Tm, const G* gq, const cplx<T>* tw,
                 const int* done, int B, const BatchStrides& bs) {
  const int M = P.M, N = P.N;
  if constexpr (sizeof(T) == 4) {
    if (M == 1024 && P.wide2_batch > 0 && B >= P.wide2_batch) {
      hipLaunchKernelGGL((k_pass2w<G>), dim3(M / kPW, B), dim3(256), 0, st, (float2*)Tm, gq, N, (const float2*)tw,
                         done, bs);
      return hipGetLastError() == hipSuccess ? CGM_SUCCESS : CGM_ERROR_DEVICE;
    }
  }
  const int cols = P.cols2;
  const int grid = M / cols;
  const int threads = cols * fft_threads(M);
  const size_t lds = (size_t)cols * lds_len(M, sizeof(T) == 8) * sizeof(cplx<T>);
  if (threads > pass_max_threads(M, sizeof(T) == 8) || lds > 65536 || M % cols) return CGM_ERROR_UNSUPPORTED;
  switch (M) {
#define X(m)                                                                                  \
  case m:                                                                                     \
    hipLaunchKernelGGL((k_pass2<T, G, m>), dim3(grid, B), dim3(threads), lds, st, Tm, gq, N,    \
                       cols, tw, done, bs);                                                   \
    break;
    CGM_FOR_EACH_M(X)
#undef X
    default: return CGM_ERROR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? CGM_SUCCESS : CGM_ERROR_DEVICE;
}

template <typename T>
int launch_pass3(const Plan& P, hipStream_t st, const cplx<T>* Tm, const cplx<T>* x, const T* chi,
                 const cplx<T>* b, cplx<T>* y, int mode, const cplx<T>* tw, const int* done,
                 int B, const BatchStrides& bs) {
  const int M = P.M, N = P.N;
  if constexpr (sizeof(T) == 4) {
    if (M == 1024 && P.wide3_batch > 0 && B >= P.wide3_batch && !b &&
        (mode == kModeForward || mode == kModeAdjoint || mode == kModeGreen)) {
      hipLaunchKernelGGL(k_pass3w, dim3((N + 7) / 8, B), dim3(256), 0, st, (const float2*)Tm, (const float2*)x,
                         (const float*)chi, (float2*)y, mode, N, (float)(1.0 / ((double)M * (double)M)),
                         (const float2*)tw, done, bs);
      return hipGetLastError() == hipSuccess ? CGM_SUCCESS : CGM_ERROR_DEVICE;
    }
  }
  const int rows = P.rows3;
  const int grid = (N + rows - 1) / rows;
  const int threads = rows * fft_threads(M);
  const size_t lds = (size_t)rows * lds_len(M, sizeof(T) == 8) * sizeof(cplx<T>);
  if (threads > pass_max_threads(M, sizeof(T) == 8) || lds > 65536) return CGM_ERROR_UNSUPPORTED;
  const T inv = T(1.0 / ((double)M * (double)M));
  switch (M) {
#define X(m)                                                                                   \
  case m:                                                                                      \
    hipLaunchKernelGGL((k_pass3<T, m>), dim3(grid, B), dim3(threads), lds, st, Tm, x, chi, b,   \
                       y, mode, N, rows, inv, tw, done, bs);                                   \
    break;
    CGM_FOR_EACH_M(X)
#undef X
    default: return CGM_ERROR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? CGM_SUCCESS : CGM_ERROR_DEVICE;
}

template int launch_pass1<float>(const Plan&, hipStream_t, const float2*, float2*, const float*,
                                 const double*, int, float2*, const float2*, const int*, int,
                                 const BatchStrides&);
template int launch_pass1<double>(const Plan&, hipStream_t, const double2*, double2*,
                                  const double*, const double*, int, double2*, const double2*,
                                  const int*, int, const BatchStrides&);
template int launch_pass2<float, float2>(const Plan&, hipStream_t, float2*, const float2*,
                                         const float2*, const int*, int, const BatchStrides&);
template int launch_pass2<double, double2>(const Plan&, hipStream_t, double2*, const double2*,
                                           const double2*, const int*, int, const BatchStrides&);
template int launch_pass3<float>(const Plan&, hipStream_t, const float2*, const float2*,
                                 const float*, const float2*, float2*, int, const float2*,
                                 const int*, int, const BatchStrides&);
template int launch_pass3<double>(const Plan&, hipStream_t, const double2*, const double2*,
                                  const double*, const double2*, double2*, int, const double2*,
                                  const int*, int, const BatchStrides&);

int plan_init(Plan& P, int N) {
  P.N = N;
  P.M = choose_embedding(N);
  if (P.M == 0) return CGM_ERROR_UNSUPPORTED;
  const int nt = fft_threads(P.M);
  P.rows1 = groups_for(P.M, 256);
  P.rows3 = P.rows1;
  P.cols2 = groups_for(P.M, 256);
  while (P.M % P.cols2) P.cols2 /= 2;
  if (P.cols2 < 1) P.cols2 = 1;
  P.wide1_batch = P.wide2_batch = P.wide3_batch = 0;
  (void)nt;
  return CGM_SUCCESS;
}

// FP64 spectrum of an N x N generator block in device memory.  Writes the
// complex64 quarter (gq32, may be null) and/or the complex128 quarter
// (gq64, may be null).
int spectrum_from_block(const Plan& P, hipStream_t st, const double2* blk, const double2* tw64,
                        float2* gq32, double2* gq64) {
  const int N = P.N, M = P.M;
  double2* E = nullptr;
  CGM_HIP(hipMallocAsync((void**)&E, sizeof(double2) * (size_t)M * M, st));
  const long MM = (long)M * M;
  hipLaunchKernelGGL(k_embed, dim3((MM + 255) / 256), dim3(256), 0, st, blk, E, N, M);
  const int nt = fft_threads(M);
  const int rows = groups_for(M, 256);
  const size_t lds = (size_t)rows * M * sizeof(double2);
  switch (M) {
#define X(m)                                                                                   \
  case m:                                                                                      \
    hipLaunchKernelGGL((k_fft_rows<double, m, -1>), dim3((M + rows - 1) / rows),                \
                       dim3(rows * nt), lds, st, E, M, rows, tw64);                             \
    hipLaunchKernelGGL((k_fft_cols<double, m, -1>), dim3(M / P.cols2), dim3(P.cols2 * nt),      \
                       (size_t)P.cols2 * M * sizeof(double2), st, E, P.cols2, tw64);            \
    break;
    CGM_FOR_EACH_M(X)
#undef X
    default:
      (void)hipFreeAsync(E, st);
      return CGM_ERROR_UNSUPPORTED;
  }
  const int H = M / 2 + 1;
  const long HH = (long)H * H;
  if (gq32) hipLaunchKernelGGL((k_quarter<float2>), dim3((HH + 255) / 256), dim3(256), 0, st, E, gq32, M);
  if (gq64) hipLaunchKernelGGL((k_quarter<double2>), dim3((HH + 255) / 256), dim3(256), 0, st, E, gq64, M);
  CGM_HIP(hipGetLastError());
  CGM_HIP(hipFreeAsync(E, st));
  return CGM_SUCCESS;
}

int green_block(hipStream_t st, double2* blk, int N, int kind, double k, double sr, double sc) {
  const long n = (long)N * N;
  hipLaunchKernelGGL(k_green_block, dim3((n + 255) / 256), dim3(256), 0, st, blk, N, kind, k, sr, sc);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

int gprep_block(hipStream_t st, const float2* gprep, double2* blk, int N) {
  const long n = (long)N * N;
  hipLaunchKernelGGL(k_gprep_block, dim3((n + 255) / 256), dim3(256), 0, st, gprep, blk, N);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

}  // namespace cgm
