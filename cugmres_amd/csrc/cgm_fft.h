// LDS-resident self-sorting Stockham FFT for one transform of compile-time
// length M, executed by NT = M / EPT threads that each hold EPT elements.
//
// Stage s with radix R over span Ns (product of earlier radices): butterfly
// j in [0, M/R) reads x[j + r*M/R], multiplies by exp(sign*2*pi*i*r*(j%Ns)/(Ns*R)),
// applies the R-point DFT and writes x[(j/Ns)*Ns*R + j%Ns + r*Ns].  Twiddles
// come from a table tw[q] = exp(-2*pi*i*q/M) built on the host in FP64.
// The unnormalised transform is the one cuFFT computes for the reference
// (B/CUDA_GMRES/Fast_matvec.cuh:13-26).
#pragma once
#include "cgm_common.h"

namespace cgm {

// ---- LDS layout ---------------------------------------------------------------
// Single-precision transforms keep one pad element after every 8 in LDS:
// element e sits at e + e/8.  The first radix-8 stage writes x[8j + r] (64-byte
// lane stride for complex64), which without the pad lands every lane of a
// wave in the same few banks; with it the stride is 72 bytes and the writes
// spread over all banks, while stride-1 reads stay contiguous.  Double
// transforms (the one-off spectrum and the FP64 refinement) are unpadded so
// that a 4096-point row still fits 64 KiB.
template <typename T> __host__ __device__ constexpr int lpad(int e) { return sizeof(T) == 4 ? e + (e >> 3) : e; }
template <typename T, int M> struct LdsLen {
  static constexpr int value = sizeof(T) == 4 ? M + M / 8 : M;
};

// ---- radix plans ------------------------------------------------------------
// Elements per thread: 8 for powers of two (radices 8/4/2), 12 for 3*2^k
// (radices 3/4/2).  Radices must divide EPT so that every thread owns exactly
// EPT/R butterflies per stage.
template <int M> struct FftShape {
  static constexpr bool pow2 = (M & (M - 1)) == 0;
  static constexpr int EPT = pow2 ? 8 : 12;
  static constexpr int NT = M / EPT;
  // work-group size bound used for launch_bounds: groups of transforms fill 256 threads
  static constexpr int WG = NT > 256 ? NT : 256;
};

template <int M> __host__ __device__ constexpr int fft_num_stages() {
  int n = M, k = 0;
  if ((M & (M - 1)) == 0) {
    while (n % 8 == 0 && n > 1) { n /= 8; ++k; }
    if (n > 1) ++k;
  } else {
    while (n % 3 == 0) { n /= 3; ++k; }
    while (n % 4 == 0) { n /= 4; ++k; }
    if (n > 1) ++k;
  }
  return k;
}

template <int M> __host__ __device__ constexpr int fft_radix(int s) {
  int n = M, k = 0;
  if ((M & (M - 1)) == 0) {
    while (n % 8 == 0 && n > 1) { if (k == s) return 8; n /= 8; ++k; }
    return n;  // remainder 2 or 4
  } else {
    while (n % 3 == 0) { if (k == s) return 3; n /= 3; ++k; }
    while (n % 4 == 0) { if (k == s) return 4; n /= 4; ++k; }
    return n;  // remainder 2
  }
}

// ---- small DFTs in registers -------------------------------------------------
template <typename T, int SIGN>
__device__ inline void dft2(cplx<T>& a, cplx<T>& b) {
  cplx<T> t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

// multiply by SIGN*i
template <typename T, int SIGN>
__device__ inline cplx<T> mul_si(cplx<T> a) {
  return SIGN < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

template <typename T, int SIGN>
__device__ inline void dft4(cplx<T>* v) {
  cplx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
  cplx<T> c = cadd(v[1], v[3]), d = mul_si<T, SIGN>(csub(v[1], v[3]));
  v[0] = cadd(a, c);
  v[2] = csub(a, c);
  v[1] = cadd(b, d);
  v[3] = csub(b, d);
}

template <typename T, int SIGN>
__device__ inline void dft3(cplx<T>* v) {
  const T c1 = T(-0.5);
  const T s1 = T(SIGN) * T(0.86602540378443864676);  // sin(2pi/3) * sign
  cplx<T> a = v[0], b = v[1], c = v[2];
  cplx<T> s = cadd(b, c), d = csub(b, c);
  v[0] = cadd(a, s);
  cplx<T> m = mk(a.x + c1 * s.x, a.y + c1 * s.y);
  // i*s1*d
  cplx<T> r = mk(-s1 * d.y, s1 * d.x);
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <typename T, int SIGN>
__device__ inline void dft8(cplx<T>* v) {
  // radix-2 split: even/odd 4-point DFTs, then combine with w8^k
  cplx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cplx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<T, SIGN>(e);
  dft4<T, SIGN>(o);
  const T h = T(0.70710678118654752440);
  // w8^1 = (1 + SIGN i)/sqrt2, w8^2 = SIGN i, w8^3 = (-1 + SIGN i)/sqrt2
  cplx<T> o1 = cmul(o[1], mk(h, T(SIGN) * h));
  cplx<T> o2 = mul_si<T, SIGN>(o[2]);
  cplx<T> o3 = cmul(o[3], mk(-h, T(SIGN) * h));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <typename T, int SIGN, int R>
__device__ inline void dft(cplx<T>* v) {
  if constexpr (R == 2) dft2<T, SIGN>(v[0], v[1]);
  else if constexpr (R == 3) dft3<T, SIGN>(v);
  else if constexpr (R == 4) dft4<T, SIGN>(v);
  else if constexpr (R == 8) dft8<T, SIGN>(v);
}

// The same DFTs when the upper half of the inputs, v[R/2 .. R-1], is zero
// (the first stage of a transform of zero-padded data): v[0 .. R/2-1] are
// read, all R outputs written.
template <typename T, int SIGN>
__device__ inline void dft4_lo(cplx<T>* v) {  // v[2] = v[3] = 0
  const cplx<T> a = v[0], d = mul_si<T, SIGN>(v[1]);
  v[2] = csub(a, v[1]);
  v[0] = cadd(a, v[1]);
  v[1] = cadd(a, d);
  v[3] = csub(a, d);
}
template <typename T, int SIGN, int R>
__device__ inline void dft_lo(cplx<T>* v) {
  if constexpr (R == 2) {
    v[1] = v[0];
  } else if constexpr (R == 4) {
    dft4_lo<T, SIGN>(v);
  } else {
    static_assert(R == 8, "half-zero inputs need an even radix");
    cplx<T> e[4] = {v[0], v[2], v[0], v[0]};  // v[4], v[6] are zero
    cplx<T> o[4] = {v[1], v[3], v[1], v[1]};  // v[5], v[7] are zero
    dft4_lo<T, SIGN>(e);
    dft4_lo<T, SIGN>(o);
    const T h = T(0.70710678118654752440);
    cplx<T> o1 = cmul(o[1], mk(h, T(SIGN) * h));
    cplx<T> o2 = mul_si<T, SIGN>(o[2]);
    cplx<T> o3 = cmul(o[3], mk(-h, T(SIGN) * h));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
}

// Twiddles of one thread for every stage with span NS > 1, flattened:
// stage S, butterfly b, r = 1..R-1 at tw_offset<M>(S) + b*(R-1) + r-1.
template <int M> __host__ __device__ constexpr int fft_stage_span(int S) {
  int ns = 1;
  for (int s = 0; s < S; ++s) ns *= fft_radix<M>(s);
  return ns;
}
template <int M> __host__ __device__ constexpr int fft_tw_offset(int S) {
  int off = 0;
  for (int s = 0; s < S; ++s)
    if (fft_stage_span<M>(s) > 1) off += (FftShape<M>::EPT / fft_radix<M>(s)) * (fft_radix<M>(s) - 1);
  return off;
}
template <int M> __host__ __device__ constexpr int fft_tw_count() {
  return fft_tw_offset<M>(fft_num_stages<M>());
}

template <typename T, int M, int S>
__device__ inline void fft_load_tw(cplx<T>* reg, const cplx<T>* __restrict__ tw, int t) {
  if constexpr (S < fft_num_stages<M>()) {
    constexpr int R = fft_radix<M>(S);
    constexpr int NS = fft_stage_span<M>(S);
    if constexpr (NS > 1) {
      constexpr int NB = FftShape<M>::EPT / R;
      constexpr int NT = FftShape<M>::NT;
      constexpr int OFF = fft_tw_offset<M>(S);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int k = (t + b * NT) % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) reg[OFF + b * (R - 1) + r - 1] = tw[(r * k) * (M / (NS * R))];
      }
    }
    fft_load_tw<T, M, S + 1>(reg, tw, t);
  }
}

// One Stockham stage for the butterflies owned by thread t.  LO: the inputs
// j + r M/R with r >= R/2 (the upper half of the data, first stage only)
// are zero and not read.  HI: only the outputs below M/2 (r < R/2 of the
// last stage) are wanted and written.
template <typename T, int M, int SIGN, int R, int NS, int OFF, bool LO = false, bool HI = false>
__device__ inline void fft_stage(cplx<T>* buf, const cplx<T>* reg, int t, bool active = true) {
  constexpr int EPT = FftShape<M>::EPT;
  constexpr int NT = FftShape<M>::NT;
  constexpr int NB = EPT / R;        // butterflies per thread
  constexpr int MR = M / R;
  // Threads outside the transform (active = false) skip the data work but
  // reach the same two barriers: the barriers stay outside any branch, so a
  // wave whose lanes are partly inside the transform (rows * NT not a multiple
  // of 64) still executes exactly two of them.
  static_assert(!LO || (NS == 1 && R % 2 == 0), "half-zero inputs: first stage, even radix");
  static_assert(!HI || (NS * R == M && R % 2 == 0), "half outputs: last stage, even radix");
  cplx<T> v[NB][R];
  if (active) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + b * NT;
#pragma unroll
      for (int r = 0; r < (LO ? R / 2 : R); ++r) v[b][r] = buf[lpad<T>(j + r * MR)];
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + b * NT;
      const int k = j % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<T> w = reg[OFF + b * (R - 1) + r - 1];
          if (SIGN > 0) w.y = -w.y;
          v[b][r] = cmul(v[b][r], w);
        }
      }
      if constexpr (LO)
        dft_lo<T, SIGN, R>(v[b]);
      else
        dft<T, SIGN, R>(v[b]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < (HI ? R / 2 : R); ++r) buf[lpad<T>(base + r * NS)] = v[b][r];
    }
  }
  __syncthreads();
}

template <typename T, int M, int SIGN, int S, int NS, bool LO, bool HI>
__device__ inline void fft_stages(cplx<T>* buf, const cplx<T>* reg, int t, bool active) {
  if constexpr (S < fft_num_stages<M>()) {
    constexpr int R = fft_radix<M>(S);
    constexpr bool lo = LO && S == 0 && R % 2 == 0;
    constexpr bool hi = HI && S == fft_num_stages<M>() - 1 && R % 2 == 0;
    fft_stage<T, M, SIGN, R, NS, fft_tw_offset<M>(S), lo, hi>(buf, reg, t, active);
    fft_stages<T, M, SIGN, S + 1, NS * R, LO, HI>(buf, reg, t, active);
  }
}

// Per-thread twiddle registers, loaded once per kernel (before the data, so
// the table reads overlap the first global loads) and reused by every
// forward and inverse transform the thread takes part in.
template <typename T, int M> struct FftTwiddles {
  static constexpr int K = fft_tw_count<M>() > 0 ? fft_tw_count<M>() : 1;
  cplx<T> reg[K];
  __device__ inline void load(const cplx<T>* __restrict__ tw, int t) { fft_load_tw<T, M, 0>(reg, tw, t); }
};

// Length of the zero padding a transform with LO still reads: the first
// stage of an even first radix never reads buf[M/2 ..], so zero-padded data
// need zeros only up to M/2 there (up to M otherwise).
template <int M> __host__ __device__ constexpr int fft_lo_read_end() { return fft_radix<M>(0) % 2 == 0 ? M / 2 : M; }

// Full transform of the M elements of buf (LDS, padded layout lpad<T>) by
// threads t in [0, NT).  The caller
// must have synchronised the work-group after filling buf; on return buf
// holds the result and the group is synchronised.  LO: buf[M/2 ..] is zero
// (zero-padded data; used when the first radix is even, the upper half is
// then never read).  HI: only the outputs below M/2 are used (when the last
// radix is even the others are not written and buf[M/2 ..] is left as it
// was).
template <typename T, int M, int SIGN, bool LO = false, bool HI = false>
__device__ inline void lds_fft(cplx<T>* buf, const FftTwiddles<T, M>& tws, int t, bool active = true) {
  fft_stages<T, M, SIGN, 0, 1, LO, HI>(buf, tws.reg, t, active);
}

}  // namespace cgm
