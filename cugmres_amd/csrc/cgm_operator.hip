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
// Register-resident 1024-point transforms, one wave per transform (complex64,
// M = 1024; the autotuner chooses them per pass and batch size).  Lane l
// holds element l + 64 q in register q (q < 16), and the transform is the
// Stockham sequence of radices 16, 16, 4:
//   stage 0: a 16-point DFT over the registers, written to LDS as x[16 l + r];
//   stage 1: reads x[l + 64 r], twiddles exp(s 2 pi i r (l % 16) / 256),
//            16-point DFT, writes x[256 (l / 16) + l % 16 + 16 r];
//   stage 2: butterflies j = l + 64 b (b < 4) read x[j + 256 r], twiddles
//            exp(s 2 pi i r j / 1024), 4-point DFT; X[j + 256 r] stays in
//            register b + 4 r, i.e. X[l + 64 q] in register q again.
// The forward output is therefore the inverse's input layout, and the column
// pass multiplies by the spectrum in registers between the two.  Two LDS
// exchanges per transform, in the wave's own region of kWL elements with one
// pad element after every 16: every access pattern above then falls in
// distinct banks within each group of 16 lanes.  A wave orders its own LDS
// accesses, so a transform needs no work-group barrier.
// ---------------------------------------------------------------------------
constexpr int kWL = 1024 + 1024 / 16;  // LDS elements per wave
__device__ inline int wpad(int e) { return e + (e >> 4); }

// Orders one wave's LDS writes before its later reads (no instruction: the
// wave's LDS accesses complete in issue order; this keeps the compiler from
// moving them across).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// a * exp(SIGN 2 pi i Q / 16)
template <int SIGN, int Q>
__device__ inline float2 rot16(float2 a) {
  constexpr float h = 0.70710678118654752440f;
  constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;  // cos, sin of pi/8
  if constexpr (Q == 4) {
    return mul_si<float, SIGN>(a);
  } else if constexpr (Q == 2) {
    return make_float2(h * (a.x - SIGN * a.y), h * (a.y + SIGN * a.x));
  } else if constexpr (Q == 6) {
    return make_float2(-h * (a.x + SIGN * a.y), h * (SIGN * a.x - a.y));
  } else if constexpr (Q == 1) {
    return cmul(a, make_float2(c1, SIGN * s1));
  } else if constexpr (Q == 3) {
    return cmul(a, make_float2(s1, SIGN * c1));
  } else {
    static_assert(Q == 9, "twiddles of the 16-point DFT");
    return cmul(a, make_float2(-c1, -SIGN * s1));
  }
}

// v[k] <- sum_n v[n] exp(SIGN 2 pi i n k / 16) with n = 4 m + p: four 4-point
// DFTs over m, the twiddles exp(SIGN 2 pi i p k1 / 16), then four 4-point DFTs
// over p give X[k1 + 4 k2].  LO: v[8 ..] are zero and not read.
template <int SIGN, bool LO>
__device__ inline void dft16(float2* v) {
  float2 a[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
#pragma unroll
    for (int m = 0; m < (LO ? 2 : 4); ++m) a[p][m] = v[4 * m + p];
    if constexpr (LO)
      dft4_lo<float, SIGN>(a[p]);
    else
      dft4<float, SIGN>(a[p]);
  }
  a[1][1] = rot16<SIGN, 1>(a[1][1]);
  a[1][2] = rot16<SIGN, 2>(a[1][2]);
  a[1][3] = rot16<SIGN, 3>(a[1][3]);
  a[2][1] = rot16<SIGN, 2>(a[2][1]);
  a[2][2] = rot16<SIGN, 4>(a[2][2]);
  a[2][3] = rot16<SIGN, 6>(a[2][3]);
  a[3][1] = rot16<SIGN, 3>(a[3][1]);
  a[3][2] = rot16<SIGN, 6>(a[3][2]);
  a[3][3] = rot16<SIGN, 9>(a[3][3]);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 b[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    dft4<float, SIGN>(b);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
  }
}

// The twiddles of lane l for stages 1 and 2 (forward sign; the inverse uses
// their conjugates), from the table tw[q] = exp(-2 pi i q / 1024).
struct WaveTw {
  float2 s1[15];  // r = 1..15: exp(-2 pi i r (l % 16) / 256)
  float2 s2[12];  // b = 0..3, r = 1..3 at 3 b + r - 1: exp(-2 pi i r (l + 64 b) / 1024)
  __device__ inline void load(const float2* __restrict__ tw, int l) {
    const int k = l & 15;
#pragma unroll
    for (int r = 1; r < 16; ++r) s1[r - 1] = tw[4 * r * k];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 1; r < 4; ++r) s2[3 * b + r - 1] = tw[r * (l + 64 * b)];
  }
};

template <int SIGN>
__device__ inline float2 twid(float2 a, float2 w) {
  if (SIGN > 0) w.y = -w.y;
  return cmul(a, w);
}

// Stages 0 and 1 of a transform of v (in the register layout above); on
// return stage 1's output is in X.  LO: v[8 ..] are zero and not read.
template <int SIGN, bool LO>
__device__ inline void wfft_head(float2* v, float2* __restrict__ X, const WaveTw& w, int l) {
  dft16<SIGN, LO>(v);
#pragma unroll
  for (int r = 0; r < 16; ++r) X[wpad(16 * l + r)] = v[r];
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = X[wpad(l + 64 * r)];
#pragma unroll
  for (int r = 1; r < 16; ++r) v[r] = twid<SIGN>(v[r], w.s1[r - 1]);
  dft16<SIGN, false>(v);
  const int base = 256 * (l >> 4) + (l & 15);
#pragma unroll
  for (int r = 0; r < 16; ++r) X[wpad(base + 16 * r)] = v[r];
  wave_lds_sync();
}

// Forward transform; v[q] = x[l + 64 q] with x[n] = 0 for n >= 512 (v[8 ..]
// not read) becomes v[q] = X[l + 64 q] for all 16 q.
__device__ inline void wfft_fwd_lo(float2* v, float2* __restrict__ X, const WaveTw& w, int l) {
  wfft_head<-1, true>(v, X, w, l);
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    float2 u[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) u[r] = X[wpad(l + 64 * b + 256 * r)];
#pragma unroll
    for (int r = 1; r < 4; ++r) u[r] = twid<-1>(u[r], w.s2[3 * b + r - 1]);
    dft4<float, -1>(u);
#pragma unroll
    for (int r = 0; r < 4; ++r) v[b + 4 * r] = u[r];
  }
  wave_lds_sync();
}

// Inverse (unnormalised) transform of v[q] = Y[l + 64 q]; only the outputs
// below 512 are formed: v[q] = y[l + 64 q] for q < 8.
__device__ inline void wfft_inv_hi(float2* v, float2* __restrict__ X, const WaveTw& w, int l) {
  wfft_head<1, false>(v, X, w, l);
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    float2 u[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) u[r] = X[wpad(l + 64 * b + 256 * r)];
#pragma unroll
    for (int r = 1; r < 4; ++r) u[r] = twid<1>(u[r], w.s2[3 * b + r - 1]);
    dft4<float, 1>(u);
    v[b] = u[0];
    v[b + 4] = u[1];
  }
  wave_lds_sync();
}

// Pass 2, M = 1024: a work-group of four waves takes half a column panel
// (N rows x 4 columns, kPW / 2 contiguous values per row), staged through LDS
// with a row stride of 5 values for coalesced global accesses; wave g then
// transforms column g of it forward, multiplies by the spectrum column and
// transforms back, keeping rows < N.
template <typename G>
__global__ void __launch_bounds__(256)
k_pass2w(float2* __restrict__ Tm, const G* __restrict__ gq, int N, const float2* __restrict__ tw,
         const int* __restrict__ done, BatchStrides bs) {
  constexpr int M = 1024;
  constexpr int H = M / 2 + 1;
  constexpr int kC = kPW / 2;  // columns per work-group
  constexpr int kRS = kC + 1;  // staging row stride
  static_assert(512 * kRS <= 4 * kWL, "the staged half panel fits the work-group's LDS");
  __shared__ float2 lds[4 * kWL];
  Tm += (long)blockIdx.y * bs.T;
  done = boff(done, (long)blockIdx.y * bs.st);
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int panel = tile >> 1;
  const int grp = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = tile * kC + grp;
  WaveTw w;
  w.load(tw, lane);
  float2* S = lds;
  float2* g0 = Tm + (long)panel * N * kPW + (tile & 1) * kC;
  for (int e = threadIdx.x; e < N * kC; e += 256) S[(e >> 2) * kRS + (e & 3)] = g0[(long)(e >> 2) * kPW + (e & 3)];
  __syncthreads();
  if (done && *done) return;  // uniform over the work-group
  float2 v[16];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int n = lane + 64 * q;
    v[q] = n < N ? S[n * kRS + grp] : make_float2(0.f, 0.f);
  }
  __syncthreads();  // the staging area is reused by the transforms
  float2* X = lds + grp * kWL;
  wfft_fwd_lo(v, X, w, lane);
  const int cq = col <= M - col ? col : M - col;
  const G* grow = gq + (long)cq * H;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int k = lane + 64 * q;
    const G g = grow[k <= M - k ? k : M - k];
    v[q] = cmul(v[q], make_float2((float)g.x, (float)g.y));
  }
  wfft_inv_hi(v, X, w, lane);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int n = lane + 64 * q;
    if (n < N) S[n * kRS + grp] = v[q];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < N * kC; e += 256) g0[(long)(e >> 2) * kPW + (e & 3)] = S[(e >> 2) * kRS + (e & 3)];
}

// Passes 1 and 3 for M = 1024 with the same transforms: one wave per row,
// four rows per work-group.  Pass 1 loads row i of chi .* x (lane l, register
// q: element l + 64 q), transforms it forward and writes T[i][l + 64 q].
// Pass 3 reads that layout, transforms it back, and lane l combines element
// l + 64 q of the row with x.  Only the plain operator modes (y = x - G(chi x),
// x - chi G(x), G x) without input scaling take these kernels.
__global__ void __launch_bounds__(256)
k_pass1w(const float2* __restrict__ x, const float* __restrict__ chi, int mul, float2* __restrict__ Tm, int N,
         const float2* __restrict__ tw, const int* __restrict__ done, BatchStrides bs) {
  __shared__ float2 lds[4 * kWL];
  {
    const long kb = blockIdx.y;
    x += kb * bs.x;
    if (chi) chi += kb * bs.chi;
    Tm += kb * bs.T;
    done = boff(done, kb * bs.st);
  }
  const int grp = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + grp;
  if (row >= N) return;  // no work-group barriers below
  const long rb = (long)row * N;
  WaveTw w;
  w.load(tw, lane);
  float2 v[16];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int n = lane + 64 * q;
    float2 e = make_float2(0.f, 0.f);
    if (n < N) {
      e = x[rb + n];
      if (mul) {
        const float c = chi[rb + n];
        e.x *= c;
        e.y *= c;
      }
    }
    v[q] = e;
  }
  if (done && *done) return;
  wfft_fwd_lo(v, lds + grp * kWL, w, lane);
#pragma unroll
  for (int q = 0; q < 16; ++q) Tm[tpos(row, lane + 64 * q, N)] = v[q];
}

__global__ void __launch_bounds__(256)
k_pass3w(const float2* __restrict__ Tm, const float2* __restrict__ x, const float* __restrict__ chi,
         float2* __restrict__ y, int mode, int N, float inv, const float2* __restrict__ tw,
         const int* __restrict__ done, BatchStrides bs) {
  __shared__ float2 lds[4 * kWL];
  {
    const long kb = blockIdx.y;
    Tm += kb * bs.T;
    x += kb * bs.x;
    if (chi) chi += kb * bs.chi;
    y += kb * bs.y;
    done = boff(done, kb * bs.st);
  }
  const int grp = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + grp;
  if (row >= N) return;  // no work-group barriers below
  WaveTw w;
  w.load(tw, lane);
  float2 v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) v[q] = Tm[tpos(row, lane + 64 * q, N)];
  if (done && *done) return;
  wfft_inv_hi(v, lds + grp * kWL, w, lane);
  const long rb = (long)row * N;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int n = lane + 64 * q;
    if (n < N) {
      const float2 g = cscale(v[q], inv);
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
      hipLaunchKernelGGL(k_pass1w, dim3((N + 3) / 4, B), dim3(256), 0, st, (const float2*)x, (const float*)chi,
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
      hipLaunchKernelGGL((k_pass2w<G>), dim3(M / (kPW / 2), B), dim3(256), 0, st, (float2*)Tm, gq, N, (const float2*)tw,
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
      hipLaunchKernelGGL(k_pass3w, dim3((N + 3) / 4, B), dim3(256), 0, st, (const float2*)Tm, (const float2*)x,
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
