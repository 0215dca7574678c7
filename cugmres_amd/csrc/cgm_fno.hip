// Fourier-neural-operator (FNO) right preconditioner: BF16 MFMA inference and
// training kernels for gfx950.
//
// The reference has no preconditioner (B/CUDA_GMRES/GMRES.cuh runs plain
// GMRES); this network is the learned extension that flexible GMRES
// (cgm_fgmres) applies to each Krylov vector.  For a unit-norm vector v on the
// N x N grid and the contrast map chi:
//
//   a   = (chi * chi_scale, N Re v, N Im v)                 per cell
//   h0  = a W_in + b_in                                      (3 -> 32 channels)
//   h_{l+1} = act_l( h_l W_l + b_l + K_l(h_l) ),  l = 0..3,  act = GELU (tanh
//             form) for l < 3, identity for l = 3
//   out = h4 W_out + b_out                                   (32 -> 2)
//   P(v) = s_id v + (out_0 + i out_1) / N                    (s_id learned)
//
// K_l is the spectral convolution of an FNO (Li et al., "Fourier neural
// operator for parametric partial differential equations", ICLR 2021) with
// retained row frequencies kx in {0..15, N-16..N-1} and column frequencies
// ky in {0..15}:
//   U[c,kx,ky]  = sum_{x,y} h[x,y,c] e^{-2 pi i (kx x + ky y)/N}
//   V[o,kx,ky]  = sum_c U[c,kx,ky] Wspec[kx,ky,c,o]
//   K(h)[x,y,o] = (1/N^2) sum_{kx,ky} w_ky Re(V[o,kx,ky] e^{2 pi i (kx x + ky y)/N}),
//                 w_0 = 1, w_ky = 2 otherwise (the inverse real FFT of the
//                 truncated spectrum).
// Only 32 column frequencies (cos and -sin of 16 ky) survive, so the column
// transforms are GEMMs against an N x 32 basis F[y,2ky] = cos(2 pi ky y/N),
// F[y,2ky+1] = -sin(2 pi ky y/N), and every layer is
//   row projection : A_x = F^T H_x          per grid row x (K = N, MFMA)
//   k_dftx         : DFT of A along x at the 32 kx (FP32)
//   k_mix          : channel mixing per mode (FP32)
//   k_invx         : inverse DFT along x (FP32) -> B_x
//   cell GEMM      : h' = act([h, F_y] [W; B_x] + b)  per cell (K = 64, MFMA)
// In the forward pass the row projection of each layer's input runs inside
// the kernel that produces it (k_lift_row, k_cell_row), for one sample the
// inverse x-DFT of each row of layers 0-2 runs in the cell GEMM of that row,
// and the last cell GEMM carries the 32 -> 2 projection (k_cell_proj):
// 1 + 3 x 3 + 4 launches for one sample, 1 + 4 x 4 for more.  h0 is a
// pointwise function of (chi, v): outside training it never reaches global
// memory, and layer 0's cell GEMM lifts its row again.  The
// backward pass has the separate kernels (k_rowproj, k_pixel) with the roles
// transposed.
//
// Activations are BF16, channel-last ([sample][x][y][32], 64 bytes per cell);
// MFMA accumulation, the spectral path and all weight gradients are FP32.
#include <math.h>

#include <vector>

#include "cgm_internal.h"
#include "cgm_reduce.h"

namespace cgm {
namespace fno {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

constexpr int C = 32;   // channels
constexpr int KY = 16;  // retained column frequencies
constexpr int KX = 32;  // retained row frequencies
constexpr int XC = 32;  // fewest rows per partial of k_dftx (sizes the partials)
constexpr int S = 36;   // LDS row stride (bf16) of k_rowproj: 16-bank offset between lane groups
constexpr int kLayerParams = C * C + C + KX * KY * C * C * 2;
constexpr long kOffLayer0 = 4 * C;
constexpr long layer_off(int l) { return kOffLayer0 + (long)l * kLayerParams; }
constexpr long kOffOut = kOffLayer0 + 4L * kLayerParams;
constexpr long kOffSkip = kOffOut + 2 * C + 2;  // weight of the identity path
constexpr long kNumParams = kOffSkip + 1;

__device__ inline bf16x8 ld8(const bf16* p) { return __builtin_bit_cast(bf16x8, *(const u32x4*)p); }
__device__ inline f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline void st4(bf16* p, float a, float b, float c, float d) {
  struct __align__(8) Q { bf16 v[4]; } q;
  q.v[0] = (bf16)a;
  q.v[1] = (bf16)b;
  q.v[2] = (bf16)c;
  q.v[3] = (bf16)d;
  *(Q*)p = q;
}

// GELU in its tanh form, 0.5 x (1 + tanh(u)), u = sqrt(2/pi) (x + 0.044715 x^3),
// evaluated as x * sigma(2u) with sigma(t) = 1 / (1 + e^-t) (the same function:
// 0.5 (1 + tanh(u)) = sigma(2u)): e^-2u = 2^(x (kE1 + kE2 x^2)) with the
// constants folded, then the hardware exp2 and reciprocal (v_exp_f32,
// v_rcp_f32, 1 ulp; the activations are stored as BF16).  Seven VALU
// instructions per activation, where an IEEE division alone takes ten; the
// cell GEMM evaluates 32 activations per cell.  Limits: x -> +inf gives
// 2^-inf = 0 and x; x -> -inf gives rcp(inf) = 0 and -0.
constexpr float kG0 = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float kG1 = 0.044715f;
constexpr float kE1 = -2.0 * 0.7978845608028654 * 1.4426950408889634;             // -2 sqrt(2/pi) log2(e)
constexpr float kE2 = -2.0 * 0.7978845608028654 * 1.4426950408889634 * 0.044715;  // kE1 * 0.044715
__device__ inline float sig2u(float x) {
  const float t = x * __builtin_fmaf(x * x, kE2, kE1);
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(t));
}
__device__ inline float gelu(float x) { return x * sig2u(x); }
// d/dx [x sigma(2u)] = sigma + x sigma (1 - sigma) 2u'
__device__ inline float dgelu(float x) {
  const float sg = sig2u(x);
  return sg + x * sg * (1.f - sg) * 2.f * kG0 * (1.f + 3.f * kG1 * x * x);
}

__device__ inline bool skip(const int* done) { return done && *done; }

__device__ const double kOne = 1.0;  // the norm of a vector that needs no scaling

// ---------------------------------------------------------------------------
// Row projection (one grid row x of one sample per iteration):
//   out[r][a][c] = sum_y Xt[a][y] Y[r][y][c]                  (Xt: 32 x N basis)
// and with WZ also the weight-gradient partials of this block
//   part[blk][a][c] += sum_y Z[r][y][a] Y[r][y][c],  part[blk][C*C + c] += sum_y Y[r][y][c].
// Wave w owns output tile (a-tile w >> 1, c-tile w & 1); Y (and Z) of the row
// are staged in LDS (row stride S) and read back as MFMA B (and A) operands.
template <bool WZ>
__global__ void __launch_bounds__(256) k_rowproj(const bf16* __restrict__ Xt, const bf16* __restrict__ Y,
                                                 const bf16* __restrict__ Z, float* __restrict__ out,
                                                 float* __restrict__ part, int N, int rows,
                                                 const int* done) {
  if (skip(done)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  bf16* ly = reinterpret_cast<bf16*>(smem_raw);
  bf16* lz = ly + (long)N * S;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int ta = w >> 1, tc = w & 1;
  f32x4 accz = {0.f, 0.f, 0.f, 0.f};
  float accb = 0.f;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    const bf16* Yr = Y + (long)r * N * C;
    const bf16* Zr = WZ ? Z + (long)r * N * C : nullptr;
    __syncthreads();
    for (int q = tid; q < N * 4; q += 256) {
      const int y = q >> 2, c8 = (q & 3) * 8;
      const u32x4 val = *(const u32x4*)(Yr + (long)y * C + c8);
      u32x2* d = reinterpret_cast<u32x2*>(ly + y * S + c8);
      d[0] = u32x2{val[0], val[1]};
      d[1] = u32x2{val[2], val[3]};
      if (WZ) {
        const u32x4 vz = *(const u32x4*)(Zr + (long)y * C + c8);
        u32x2* dz = reinterpret_cast<u32x2*>(lz + y * S + c8);
        dz[0] = u32x2{vz[0], vz[1]};
        dz[1] = u32x2{vz[2], vz[3]};
      }
    }
    __syncthreads();
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int y0 = 0; y0 < N; y0 += 32) {
      const bf16x8 a = ld8(Xt + (long)(16 * ta + i) * N + y0 + 8 * g);
      bf16x8 b;
#pragma unroll
      for (int v = 0; v < 8; ++v) b[v] = ly[(y0 + 8 * g + v) * S + 16 * tc + i];
      acc = mma(a, b, acc);
      if (WZ) {
        bf16x8 az;
#pragma unroll
        for (int v = 0; v < 8; ++v) az[v] = lz[(y0 + 8 * g + v) * S + 16 * ta + i];
        accz = mma(az, b, accz);
      }
    }
    float* o = out + (long)r * C * C;
#pragma unroll
    for (int v = 0; v < 4; ++v) o[(16 * ta + 4 * g + v) * C + 16 * tc + i] = acc[v];
    if (WZ)
      for (int y = tid >> 5; y < N; y += 8) accb += (float)ly[y * S + (tid & 31)];
  }
  if (WZ) {
    float* pp = part + (long)blockIdx.x * (C * C + C);
#pragma unroll
    for (int v = 0; v < 4; ++v) pp[(16 * ta + 4 * g + v) * C + 16 * tc + i] = accz[v];
    __shared__ float sb[256];
    sb[tid] = accb;
    __syncthreads();
    if (tid < C) {
      float sacc = 0.f;
      for (int k = 0; k < 8; ++k) sacc += sb[tid + 32 * k];
      pp[C * C + tid] = sacc;
    }
  }
}

// ---------------------------------------------------------------------------
// DFT along x at the retained kx, one chunk of XR rows per block:
//   out[b][chunk][kxi][ky][c] = s_ky sum_{x in chunk} (in[b][x][2ky][c] + i in[b][x][2ky+1][c]) e^{-2 pi i kx x/N}
// with s_ky = 1 (forward) or w_ky / N^2 (backward of the inverse transform).
// A chunk is XR/2 row pairs: pair p >= 1 is rows (p, N - p), whose twiddles
// are conjugate (kx integer), so with P = A_x + A_{N-x}, M = A_x - A_{N-x}
// and e = e^{2 pi i kx x/N} = c + is the pair adds c P - i s M (half the
// products of two rows); pair 0 is rows 0 and N/2 (twiddles 1 and
// (-1)^kx = (-1)^kxi: P at even kxi, M at odd).  The partials of all chunks
// sum to the whole transform.  Thread (kp, cp) of 256 accumulates the 2 x 2
// outputs kxi in {2kp, 2kp+1}, c in {cp, cp + 16}.
template <int XR>
__global__ void __launch_bounds__(256) k_dftx(const float* __restrict__ in, float2* __restrict__ outp,
                                              const float2* __restrict__ Ex, int N, int scaled,
                                              const int* done) {
  if (skip(done)) return;
  constexpr int XH = XR / 2;  // pairs per chunk
  __shared__ float2 sp[XH][C + 1];
  __shared__ float2 sm[XH][C + 1];
  __shared__ float2 se[KX][XH + 1];
  const int chunk = blockIdx.x, ky = blockIdx.y, b = blockIdx.z;
  const int nch = gridDim.x;
  const int tid = threadIdx.x;
  const int p0 = chunk * XH;
  // every load of the staging first, then the LDS writes (a load-wait-write
  // chain per item would expose the load latency once per item)
  constexpr int NA = XH * C / 256, NE = KX * XH / 256;
  static_assert(NA * 256 == XH * C && NE * 256 == KX * XH, "staging items must fill the work-group");
  float2 u1[NA], u2[NA], ev[NE];
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const int e = tid + 256 * k, pl = e >> 5, c = e & 31;
    const int x = p0 + pl, x2 = x == 0 ? N / 2 : N - x;
    const float* r1 = in + ((long)(b * N + x) * C + 2 * ky) * C;
    const float* r2 = in + ((long)(b * N + x2) * C + 2 * ky) * C;
    u1[k] = make_float2(r1[c], r1[C + c]);
    u2[k] = make_float2(r2[c], r2[C + c]);
  }
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int e = tid + 256 * k, q = e / XH, pl = e % XH;
    ev[k] = Ex[(long)q * N + p0 + pl];  // frequency q at row p0 + pl
  }
#pragma unroll
  for (int k = 0; k < NA; ++k) {
    const int e = tid + 256 * k, pl = e >> 5, c = e & 31;
    sp[pl][c] = make_float2(u1[k].x + u2[k].x, u1[k].y + u2[k].y);
    sm[pl][c] = make_float2(u1[k].x - u2[k].x, u1[k].y - u2[k].y);
  }
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int e = tid + 256 * k;
    se[e / XH][e % XH] = ev[k];
  }
  __syncthreads();
  const int kp = tid >> 4, cp = tid & 15;
  float2 acc[2][2] = {};
  int pl0 = 0;
  if (p0 == 0) {  // rows 0 and N/2: P at kxi = 2kp, M at 2kp + 1
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      acc[0][c] = sp[0][cp + 16 * c];
      acc[1][c] = sm[0][cp + 16 * c];
    }
    pl0 = 1;
  }
#pragma unroll 4
  for (int pl = pl0; pl < XH; ++pl) {
    const float2 e[2] = {se[2 * kp][pl], se[2 * kp + 1][pl]};
    const float2 P[2] = {sp[pl][cp], sp[pl][cp + 16]};
    const float2 M[2] = {sm[pl][cp], sm[pl][cp + 16]};
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        acc[a][c].x += e[a].x * P[c].x + e[a].y * M[c].y;  // c P - i s M
        acc[a][c].y += e[a].x * P[c].y - e[a].y * M[c].x;
      }
  }
  float sc = 1.f;
  if (scaled) sc = (ky == 0 ? 1.f : 2.f) / ((float)N * (float)N);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
      outp[(((long)(b * nch + chunk) * KX + 2 * kp + a) * KY + ky) * C + cp + 16 * c] =
          make_float2(acc[a][c].x * sc, acc[a][c].y * sc);
}

// ---------------------------------------------------------------------------
// Channel mixing per mode, one (ky, sample) per block:
//   U[kxi][ch] = sum of the k_dftx partials             (saved to usave)
//   forward : V[kxi][o] = s_ky sum_c U[kxi][c] W[kxi][ky][c][o],  s_ky = w_ky / N^2
//   backward: V[kxi][c] = sum_o conj(W[kxi][ky][c][o]) U[kxi][o]
// V goes to vout[b][kxi][ky][ch].
constexpr int KXG = 8;  // retained kx per block of k_mix
template <bool BWD>
__global__ void __launch_bounds__(256) k_mix(const float2* __restrict__ upart, int nch,
                                             const float2* __restrict__ W, float2* __restrict__ usave,
                                             float2* __restrict__ vout, int N, const int* done) {
  if (skip(done)) return;
  __shared__ float2 su[KXG][C + 1];
  const int ky = blockIdx.x, b = blockIdx.z;
  const int tid = threadIdx.x;
  const int kl = tid >> 5, ch = tid & 31;
  const int kxi = blockIdx.y * KXG + kl;
  float2 acc = make_float2(0.f, 0.f);
  for (int k0 = 0; k0 < nch; k0 += 8) {
    float2 u[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      u[j] = (k0 + j < nch) ? upart[(((long)(b * nch + k0 + j) * KX + kxi) * KY + ky) * C + ch]
                            : make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      acc.x += u[j].x;
      acc.y += u[j].y;
    }
  }
  su[kl][ch] = acc;
  const long slot = (((long)b * KX + kxi) * KY + ky) * C + ch;
  if (usave) usave[slot] = acc;
  __syncthreads();
  const float2* Wm = W + (long)(kxi * KY + ky) * C * C;
  float2 v = make_float2(0.f, 0.f);
  if (!BWD) {
    // all 32 weights of this output in flight at once (8 at a time would
    // expose the load latency four times)
    float2 wv[C];
#pragma unroll
    for (int c = 0; c < C; ++c) wv[c] = Wm[c * C + ch];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const float2 uu = su[kl][c];
      v.x += uu.x * wv[c].x - uu.y * wv[c].y;
      v.y += uu.x * wv[c].y + uu.y * wv[c].x;
    }
    const float sc = (ky == 0 ? 1.f : 2.f) / ((float)N * (float)N);
    v.x *= sc;
    v.y *= sc;
  } else {
#pragma unroll 8
    for (int o = 0; o < C; ++o) {
      const float2 uu = su[kl][o], wv = Wm[ch * C + o];
      v.x += wv.x * uu.x + wv.y * uu.y;  // conj(w) * u
      v.y += wv.x * uu.y - wv.y * uu.x;
    }
  }
  vout[slot] = v;
}

// Inverse DFT along x, XP row pairs of one sample per block (thread = (ky, ch)):
//   out[b][x][ch][2ky], [2ky+1] = Re, Im of sum_kxi V[kxi][ky][ch] e^{+2 pi i kx x/N}   (BF16)
// Rows x and N - x have conjugate twiddles (kx integer), so with V = a + ib,
// e = c + is and the sums S1 = sum a c, S2 = sum b s, S3 = sum a s,
// S4 = sum b c (four products per mode for two rows):
//   B_x = (S1 - S2) + i (S3 + S4),  B_{N-x} = (S1 + S2) + i (S4 - S3).
// Pair p >= 1 is rows (p, N - p); pair 0 is rows 0 and N/2, whose twiddles
// are 1 and (-1)^kx = (-1)^kxi (N even, kx = kxi or N - 32 + kxi).  The
// block's 2 XP x 32 x 32 outputs are assembled in LDS and leave as 16-byte
// stores of whole 2 KiB rows.  Every block reads its sample's 128 KiB of
// modes: XP = 8 from 8 vectors and XP = 4 from 4 (fewer reads), XP = 1 for
// fewer (enough blocks).
template <int XP>
__global__ void __launch_bounds__(512) k_invx(const float2* __restrict__ V, bf16* __restrict__ out,
                                              const float2* __restrict__ Ex, int N, const int* done) {
  if (skip(done)) return;
  __shared__ float2 se[XP][KX];
  __shared__ __align__(16) bf16 lt[2 * XP][C][C];
  const int p0 = blockIdx.x * XP, b = blockIdx.z;
  const int tid = threadIdx.x;
  if (tid < XP * KX) {
    const int q = tid / XP, pl = tid % XP;
    se[pl][q] = Ex[(long)q * N + p0 + pl];
  }
  const int ch = tid & 31, ky = tid >> 5;
  float2 vq[KX];
#pragma unroll
  for (int q = 0; q < KX; ++q) vq[q] = V[(((long)b * KX + q) * KY + ky) * C + ch];
  __syncthreads();
#pragma unroll 2
  for (int pl = 0; pl < XP; ++pl) {
    float s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
#pragma unroll
    for (int q = 0; q < KX; ++q) {
      const float2 e = se[pl][q];
      s1 += vq[q].x * e.x;
      s2 += vq[q].y * e.y;
      s3 += vq[q].x * e.y;
      s4 += vq[q].y * e.x;
    }
    lt[pl][ch][2 * ky] = (bf16)(s1 - s2);
    lt[pl][ch][2 * ky + 1] = (bf16)(s3 + s4);
    if (p0 + pl > 0) {
      lt[XP + pl][ch][2 * ky] = (bf16)(s1 + s2);
      lt[XP + pl][ch][2 * ky + 1] = (bf16)(s4 - s3);
    } else {  // row N/2
      float t1 = 0.f, t2 = 0.f;
#pragma unroll
      for (int q = 0; q < KX; q += 2) {
        t1 += vq[q].x - vq[q + 1].x;
        t2 += vq[q].y - vq[q + 1].y;
      }
      lt[XP][ch][2 * ky] = (bf16)t1;
      lt[XP][ch][2 * ky + 1] = (bf16)t2;
    }
  }
  __syncthreads();
  // staged row r: x = p0 + r (r < XP), else N - (p0 + r - XP), pair 0's N/2
  const u32x4* src = reinterpret_cast<const u32x4*>(&lt[0][0][0]);
  for (int q = tid; q < 2 * XP * C * C / 8; q += 512) {
    const int r = q / (C * C / 8), pl = r < XP ? r : r - XP;
    const int x = r < XP ? p0 + pl : (p0 + pl == 0 ? N / 2 : N - p0 - pl);
    reinterpret_cast<u32x4*>(out + ((long)b * N + x) * C * C)[q - r * (C * C / 8)] = src[q];
  }
}

// Spectral weight gradient: G[kxi][ky][c][o] = sum_b conj(U[b][kxi][ky][c]) dV[b][kxi][ky][o].
__global__ void __launch_bounds__(256) k_specgrad(const float2* __restrict__ U, const float2* __restrict__ dV,
                                                  float2* __restrict__ G, int B) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)KX * KY * C * C) return;
  const int o = (int)(e % C);
  const int c = (int)((e / C) % C);
  const long mode = e / (C * C);  // kxi * KY + ky
  float2 acc = make_float2(0.f, 0.f);
  for (int b = 0; b < B; ++b) {
    const float2 u = U[((long)b * KX * KY + mode) * C + c];
    const float2 d = dV[((long)b * KX * KY + mode) * C + o];
    acc.x += u.x * d.x + u.y * d.y;
    acc.y += u.x * d.y - u.y * d.x;
  }
  G[e] = acc;
}

// ---------------------------------------------------------------------------
// Per-cell GEMM with K = 64 and fused epilogue:
//   out[p][o] = epi( sum_c P[p][c] A[o][c] + sum_k Q[y][k] R[b][x][o][k] + bias[o] )
// EPI 0: identity, 1: GELU (pre-activation to pre_out when given),
// 2: times GELU'(aux[p][o]).  The product is formed transposed (rows = output
// channels, columns = 16 cells) so that lane l holds 4 consecutive channels of
// cell l % 16 in both the B operand and the accumulator.  One wave covers 64
// consecutive cells of a grid row.
template <int EPI>
__global__ void __launch_bounds__(256) k_pixel(const bf16* __restrict__ P, const bf16* __restrict__ Q,
                                               const bf16* __restrict__ A, const bf16* __restrict__ R,
                                               const float* __restrict__ bias, bf16* __restrict__ pre_out,
                                               const bf16* __restrict__ aux, bf16* __restrict__ out, int N,
                                               const int* done) {
  if (skip(done)) return;
  // each wave stages its 16-cell tile (16 x 32 BF16, rows padded to 80 B) so
  // that every lane stores 16 contiguous bytes and a wave 1 KiB at once
  __shared__ __align__(16) bf16 stage[2][4][16][40];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long wv = (long)blockIdx.x * 4 + w;  // every block is full: waves % 4 == 0
  const int segs = N / 64;
  const long bx = wv / segs;  // sample * N + x
  const int y0 = (int)(wv % segs) * 64;
  const int i = lane & 15, g = lane >> 4;
  const int rrow = lane >> 2, rcol = (lane & 3) * 8;
  bf16x8 aw[2], ar[2];
  float bs[2][4];
#pragma unroll
  for (int to = 0; to < 2; ++to) {
    aw[to] = ld8(A + (16 * to + i) * C + 8 * g);
    ar[to] = ld8(R + (bx * C + 16 * to + i) * C + 8 * g);
#pragma unroll
    for (int v = 0; v < 4; ++v) bs[to][v] = bias ? bias[16 * to + 4 * g + v] : 0.f;
  }
  // all four tiles' operands in flight before the first MFMA
  bf16x8 bpt[4], bqt[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int y = y0 + 16 * t + i;
    bpt[t] = ld8(P + (bx * N + y) * C + 8 * g);
    bqt[t] = ld8(Q + (long)y * C + 8 * g);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int y = y0 + 16 * t + i;
    const long p = bx * N + y;
    const bf16x8 bp = bpt[t];
    const bf16x8 bq = bqt[t];
#pragma unroll
    for (int to = 0; to < 2; ++to) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mma(aw[to], bp, acc);
      acc = mma(ar[to], bq, acc);
      const int o0 = 16 * to + 4 * g;
      float r[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) r[v] = acc[v] + bs[to][v];
      if (EPI == 1) {
        if (pre_out) st4(&stage[1][w][i][o0], r[0], r[1], r[2], r[3]);
#pragma unroll
        for (int v = 0; v < 4; ++v) r[v] = gelu(r[v]);
      } else if (EPI == 2) {
#pragma unroll
        for (int v = 0; v < 4; ++v) r[v] *= dgelu((float)aux[p * C + o0 + v]);
      }
      st4(&stage[0][w][i][o0], r[0], r[1], r[2], r[3]);
    }
    __syncthreads();
    const long po = (bx * N + y0 + 16 * t + rrow) * C + rcol;
    *(u32x4*)(out + po) = *(const u32x4*)&stage[0][w][rrow][rcol];
    if (EPI == 1 && pre_out) *(u32x4*)(pre_out + po) = *(const u32x4*)&stage[1][w][rrow][rcol];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Forward-pass kernels that own whole grid rows, so that the row projection of
// the activations they produce (the first step of the next layer's spectral
// convolution) runs on the row while it is still in LDS:
//   k_lift_row  : lift of one row, then A_x = F^T H0_x
//   k_cell_row  : B_x of the row, cell GEMM + GELU of the row (layers 0-2),
//                 then A_x = F^T H'_x
//   k_cell_proj : cell GEMM of layer 3 with the 32 -> 2 projection and the
//                 identity path, z = s_id v + (h4 W_out + b_out) / s
// The row kernels stage the row they project channel-major, [c][y] BF16 with
// row stride basis_ld(N) (the K = y operand of the projection is then 16
// contiguous bytes per lane), and the projection basis F^T likewise.

// Row stride (BF16) of the channel-major row and of the projection basis F^T
// staged in LDS by the row kernels: 16 bytes of padding move row a + 1 four
// banks on, so the 16 rows an MFMA operand read touches fall in different
// banks.
__host__ __device__ constexpr int basis_ld(int N) { return N + 8; }
// LDS bytes of the forward row kernels: the basis and the channel-major row
// or (layer 0's cell GEMM) the row as [y][S].
inline size_t row_kernel_lds(int N) {
  const size_t t = (size_t)C * basis_ld(N) * sizeof(bf16), r = (size_t)N * S * sizeof(bf16);
  return t + (t > r ? t : r);
}

// F^T (32 x N, global) to LDS with row stride basis_ld(N), 16 bytes per item,
// four items per thread in flight together (the loads are unconditional, on
// a clamped item, so that no branch waits for one load before the next is
// issued).
template <int NT>
__device__ inline void basis_to_lds(const bf16* __restrict__ Xt, bf16* lb, int N) {
  const int per_row = N / 8, items = C * per_row;
  u32x4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int q = threadIdx.x + k * NT;
    q = q < items ? q : items - 1;
    const int a = q / per_row, y8 = (q - a * per_row) * 8;
    v[k] = *reinterpret_cast<const u32x4*>(Xt + (long)a * N + y8);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int q = threadIdx.x + k * NT;
    if (q < items) {
      const int a = q / per_row, y8 = (q - a * per_row) * 8;
      *reinterpret_cast<u32x4*>(lb + a * basis_ld(N) + y8) = v[k];
    }
  }
  for (int q = threadIdx.x + 4 * NT; q < items; q += NT) {  // N > 16 NT / C only
    const int a = q / per_row, y8 = (q - a * per_row) * 8;
    *reinterpret_cast<u32x4*>(lb + a * basis_ld(N) + y8) = *reinterpret_cast<const u32x4*>(Xt + (long)a * N + y8);
  }
}

// A_x = F^T H_x from the row staged channel-major in LDS (lt[c][y], row
// stride basis_ld(N): the B operand is one 16-byte read), F^T read with row
// stride ldx (from global memory or LDS).  NW waves: wave w owns output tile
// (a-tile (w >> 1) & 1, c-tile w & 1); with NW = 8 waves w and w + 4 split K
// in halves and the upper half adds into the lower through red.
template <int NW>
__device__ inline void row_project(const bf16* __restrict__ Xt, int ldx, const bf16* lt, int N,
                                   float* __restrict__ o, float* red) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int ta = (w >> 1) & 1, tc = w & 1;
  const int kh = NW == 8 ? (w >> 2) : 0;
  const int ylo = NW == 8 ? kh * (N / 2) : 0;
  const int yhi = NW == 8 ? ylo + N / 2 : N;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  // K steps of 32 in groups of 4 with the basis loads of a group in flight
  // together (the basis rows come from L2; one load per step would expose
  // its latency N/32 times)
  const int steps = (yhi - ylo) / 32;
  const bf16* xa = Xt + (long)(16 * ta + i) * ldx + ylo + 8 * g;
  const bf16* xb = lt + (16 * tc + i) * basis_ld(N) + ylo + 8 * g;
  for (int s0 = 0; s0 < steps; s0 += 4) {
    bf16x8 a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (s0 + u < steps) a[u] = ld8(xa + 32 * (s0 + u));
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (s0 + u < steps) b[u] = ld8(xb + 32 * (s0 + u));
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (s0 + u < steps) acc = mma(a[u], b[u], acc);
  }
  if (NW == 8) {
    if (kh == 1) {
#pragma unroll
      for (int v = 0; v < 4; ++v) red[((w & 3) * 64 + lane) * 4 + v] = acc[v];
    }
    __syncthreads();
    if (kh == 1) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] += red[(w * 64 + lane) * 4 + v];
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) o[(16 * ta + 4 * g + v) * C + 16 * tc + i] = acc[v];
}

// Input of the lift: sample k reads chi at k * sChi, v at k * sV and its
// norm at nrm + k * sNrm bytes (nrm null: v is used as it is).
struct LiftIn {
  const float* chi;
  long sChi;
  float chi_scale;
  const float2* v;
  long sV;
  const double* nrm;
  long sNrm;
  float s;          // v's scale (N: unit vectors become O(1) per cell)
  const float* P;   // W_in (3 x 32) followed by b_in (32)
};

// h0 of one cell at 8 channels c0..c0+7 from a = (chi chi_scale, s Re v',
// s Im v') and the lift parameters w (W_in rows 0-2 and b_in, channels
// c0..c0+7 at w[0..7], w[8..15], w[16..23], w[24..31]):
//   h0[c] = a0 W_in[0][c] + a1 W_in[1][c] + a2 W_in[2][c] + b_in[c]
// The one rounding of the lift: both row kernels that lift call this.
__device__ inline bf16x8 lift8(float a0, float a1, float a2, const float* w) {
  bf16x8 h;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    h[k] = (bf16)__builtin_fmaf(a2, w[16 + k], __builtin_fmaf(a1, w[8 + k], __builtin_fmaf(a0, w[k], w[24 + k])));
  return h;
}

// The lift of grid row xr of sample ks into the LDS row lt, channel-major
// [c][y] with row stride basis_ld(N) (TR) or [y][S], with (a0, a1, a2) =
// (chi chi_scale, s Re v', s Im v'), v' = v / *nrm: one item q = (cell y,
// quarter) of 8 channels per thread and step (the four items of a cell are
// adjacent lanes of one load, and a thread's loads all precede its stores,
// so an in-place vw == v is read before it is written).
// With h0 the row also goes to global memory (16-byte stores), with vw v'
// goes back to vw.  In two halves: lift_load issues the loads of one group
// of four items per thread (items q0 + u NT, clamped to the row, and the
// sample's norm, all unconditional), lift_store forms and stores them.
struct LiftRegs {
  float cv[4];
  float2 xv[4];
  double d;  // the sample's norm (1 without nrm)
};
template <int NT>
__device__ inline void lift_load(const LiftIn& L, long ks, int xr, int N, int q0, LiftRegs& g) {
  const long rin = (long)xr * N;  // the row's first cell in the sample
  const float* chi = L.chi + ks * L.sChi;
  const float2* v = L.v + ks * L.sV;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int q = q0 + u * NT < N * 4 ? q0 + u * NT : N * 4 - 1;
    g.cv[u] = chi[rin + (q >> 2)];
    g.xv[u] = v[rin + (q >> 2)];
  }
  // a global (not flat) load, which LDS waits do not wait for
  typedef __attribute__((address_space(1))) const double gdouble;
  g.d = *(gdouble*)(L.nrm ? boff(L.nrm, ks * L.sNrm) : &kOne);
}
// The 4 x 8 lift parameters of this thread's channels (NT % 4 == 0: every
// item of a thread has the same 8 channels).
__device__ inline void lift_weights(const LiftIn& L, float* w) {
  const int c8 = (threadIdx.x & 3) * 8;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
#pragma unroll
    for (int j = 0; j < 4; ++j) w[8 * j + k] = L.P[j * C + c8 + k];
  }
}
template <int NT, bool TR>
__device__ inline void lift_store(const LiftIn& L, const LiftRegs& g, const float* w, long ks, int xr, int N,
                                  int q0, bf16* lt, bf16* h0, float2* vw) {
  static_assert(NT % 4 == 0, "items of a thread must share their channels");
  const long rin = (long)xr * N;
  const int c8 = (threadIdx.x & 3) * 8;
  const float r = L.nrm ? (g.d > 0.0 ? (float)(1.0 / g.d) : 0.f) : 1.f;
  if (vw) vw += ks * L.sV;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int q = q0 + u * NT;
    if (q >= N * 4) break;
    const int y = q >> 2;
    const long pin = rin + y;
    float2 vv = g.xv[u];
    if (L.nrm) {
      vv.x *= r;
      vv.y *= r;
      if (vw && c8 == 0) vw[pin] = vv;
    }
    const bf16x8 hq = lift8(g.cv[u] * L.chi_scale, vv.x * L.s, vv.y * L.s, w);
    if (h0) *reinterpret_cast<u32x4*>(h0 + ((ks * N + xr) * N + y) * C + c8) = __builtin_bit_cast(u32x4, hq);
    if (TR) {
#pragma unroll
      for (int k = 0; k < 8; ++k) lt[(c8 + k) * basis_ld(N) + y] = hq[k];
    } else {
      const u32x4 bits = __builtin_bit_cast(u32x4, hq);
      u32x2* d = reinterpret_cast<u32x2*>(lt + y * S + c8);
      d[0] = u32x2{bits[0], bits[1]};
      d[1] = u32x2{bits[2], bits[3]};
    }
  }
}
template <int NT, bool TR>
__device__ inline void lift_row_to_lds(const LiftIn& L, long ks, int xr, int N, bf16* lt, bf16* h0,
                                       float2* vw) {
  float w[32];
  lift_weights(L, w);
  for (int q0 = threadIdx.x; q0 < N * 4; q0 += 4 * NT) {
    LiftRegs g;
    lift_load<NT>(L, ks, xr, N, q0, g);
    lift_store<NT, TR>(L, g, w, ks, xr, N, q0, lt, h0, vw);
  }
}

// One block (512 threads) per grid row x of one sample: the lift of the row
// to LDS (and to h0 when given: training keeps it for the backward pass),
// then its row projection into Aout[x].  With nrm the input is v / *nrm
// (written back to vw when given), which lets FGMRES hand over its lazily
// normalised Krylov vector.
__global__ void __launch_bounds__(512) k_lift_row(LiftIn L, float2* vw, bf16* __restrict__ h0,
                                                  const bf16* __restrict__ Xt, float* __restrict__ Aout, int N,
                                                  const int* done) {
  if (skip(done)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  bf16* lt = reinterpret_cast<bf16*>(smem_raw);  // the row, channel-major
  bf16* lbas = lt + C * basis_ld(N);
  __shared__ float red[4 * 64 * 4];
  const long bx = blockIdx.x;
  const long ks = bx / N;  // sample
  basis_to_lds<512>(Xt, lbas, N);
  lift_row_to_lds<512, true>(L, ks, (int)(bx - ks * N), N, lt, h0, vw);
  __syncthreads();
  row_project<8>(lbas, basis_ld(N), lt, N, Aout + bx * C * C, red);
}

// k_lift_row with RPW grid rows of one sample per block (N % RPW == 0 and
// N <= 512, one group of four items per thread): the basis is staged once,
// and each next row's chi and v are loaded before the current row's
// projection.
template <int RPW>
__global__ void __launch_bounds__(512) k_lift_rows(LiftIn L, float2* vw, bf16* __restrict__ h0,
                                                   const bf16* __restrict__ Xt, float* __restrict__ Aout, int N,
                                                   const int* done) {
  if (skip(done)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  bf16* lt = reinterpret_cast<bf16*>(smem_raw);  // the row, channel-major
  bf16* lbas = lt + C * basis_ld(N);
  __shared__ float red[4 * 64 * 4];
  const long bx0 = (long)blockIdx.x * RPW;
  const long ks = bx0 / N;  // sample of all RPW rows
  const int xr0 = (int)(bx0 - ks * N);
  LiftRegs g;
  lift_load<512>(L, ks, xr0, N, threadIdx.x, g);
  basis_to_lds<512>(Xt, lbas, N);
#pragma unroll
  for (int rr = 0; rr < RPW; ++rr) {
    if (rr) __syncthreads();  // the previous row's projection has read lt
    {
      float w[32];
      lift_weights(L, w);
      lift_store<512, true>(L, g, w, ks, xr0 + rr, N, threadIdx.x, lt, h0, vw);
    }
    __syncthreads();
    if (rr + 1 < RPW) lift_load<512>(L, ks, xr0 + rr + 1, N, threadIdx.x, g);
    row_project<8>(lbas, basis_ld(N), lt, N, Aout + (bx0 + rr) * C * C, red);
  }
}

// One block (8 waves) per grid row x of sample k (layers 0-2):
//   1. B_x to LDS (BF16), the MFMA A operand of the spectral half of the
//      GEMM: copied from R[k][x] (k_invx's output), or with FOLD formed
//      here as the inverse x-DFT of the mixed modes at this row (thread
//      (ky, o): B_x[o][2ky] + i B_x[o][2ky+1] = sum_kxi Vm[k][kxi][ky][o]
//      e^{+2 pi i kx x/N}), which saves k_invx's launch but reads all of
//      Vm[k] (128 KiB, from L2) per row: the better choice for one sample;
//   2. h'[y] = GELU([h[y], F_y] [W; B_x] + b) for the row's N cells (16-cell
//      tiles t = wave + 8u, as in k_pixel) to out (each lane stores the 4
//      channels it holds of one cell, 8 bytes) and channel-major into LDS
//      (pre-activations to pre_out when given, for training);
//   3. the row projection of h' into Anext[x], with the basis staged in LDS.
// With LIFT (layer 0) h[y] is not read from P: the row is lifted from (chi,
// v) into LDS as [y][S] first (lift_row_to_lds, the rounding of k_lift_row),
// 12 bytes per cell instead of 64, where the basis goes otherwise; the tiles
// read their cells from there, and the basis replaces the row after the
// tile loop.  (Lifting in registers per tile instead holds the 32 lift
// weights of a lane through the tile loop: 170 VGPRs, one work-group per CU;
// so does holding the basis in registers through the loop: 142 VGPRs.)
template <bool FOLD, bool LIFT>
__global__ void __launch_bounds__(512) k_cell_row(const bf16* __restrict__ P, LiftIn L, const bf16* __restrict__ Q,
                                                  const bf16* __restrict__ A, const bf16* __restrict__ R,
                                                  const float2* __restrict__ Vm, const float2* __restrict__ Ex,
                                                  const float* __restrict__ bias, bf16* __restrict__ pre_out,
                                                  bf16* __restrict__ out, const bf16* __restrict__ Xt,
                                                  float* __restrict__ Anext, int N, const int* done) {
  if (skip(done)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // region 1: the basis F^T (row stride basis_ld(N)); with LIFT first the
  // lifted row h0 as [y][S], the basis after the tiles have read it
  bf16* lbas = reinterpret_cast<bf16*>(smem_raw);
  bf16* lrow = lbas;
  bf16* lt = lbas + (N * S > C * basis_ld(N) ? N * S : C * basis_ld(N));  // region 2: h' channel-major
  __shared__ float red[4 * 64 * 4];
  __shared__ __align__(16) bf16 lb[C][C + 8];  // B_x as [o][k], rows padded to 80 B
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long bx = blockIdx.x;  // sample * N + x
  const long ks = bx / N;
  const int xr = (int)(bx - ks * N);
  const int i = lane & 15, g = lane >> 4;
  const int ntile = N / 16;
  // the first tiles' cell operands and B_x are in flight while the basis
  // (or the lifted row) is staged; the loads are unconditional (tiles past
  // the row's end re-read its last tile, unused), so that no branch waits for
  // one load before the next is issued
  // the layer's weights and bias: loaded first, except with FOLD, whose
  // inverse x-DFT holds 64 registers of modes until B_x is formed
  bf16x8 aw[2];
  float bs[2][4];
  auto load_weights = [&]() {
#pragma unroll
    for (int to = 0; to < 2; ++to) {
      aw[to] = ld8(A + (16 * to + i) * C + 8 * g);
#pragma unroll
      for (int v = 0; v < 4; ++v) bs[to][v] = bias[16 * to + 4 * g + v];
    }
  };
  if (!FOLD) load_weights();
  bf16x8 bpt[4], bqt[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    int t = w + 8 * u;
    t = t < ntile ? t : ntile - 1;
    const int y = 16 * t + i;
    if (!LIFT) bpt[u] = ld8(P + (bx * N + y) * C + 8 * g);
    bqt[u] = ld8(Q + (long)y * C + 8 * g);
  }
  u32x4 bxv;  // 128 threads x 8 BF16 of B_x (the other threads read a copy)
  if (!FOLD) bxv = *reinterpret_cast<const u32x4*>(R + (bx * C + ((tid >> 2) & 31)) * C + (tid & 3) * 8);
  if (LIFT)
    lift_row_to_lds<512, false>(L, ks, xr, N, lrow, nullptr, nullptr);
  else
    basis_to_lds<512>(Xt, lbas, N);
  if (!FOLD) {
    if (tid < C * C / 8) *reinterpret_cast<u32x4*>(&lb[tid >> 2][(tid & 3) * 8]) = bxv;
  } else {
    // the sums and their order of k_invx (rows x and N - x from the
    // twiddles of the pair's first row p), so one sample gets the B_x that
    // k_invx forms for a batch, bit for bit
    const int ky = tid >> 5, o = tid & 31;
    const float2* Vs = Vm + (ks * KX * KY + ky) * C + o;
    float2 vq[KX];
#pragma unroll
    for (int q = 0; q < KX; ++q) vq[q] = Vs[(long)q * KY * C];
    float bre, bim;
    if (xr == N / 2) {
      float t1 = 0.f, t2 = 0.f;
#pragma unroll
      for (int q = 0; q < KX; q += 2) {
        t1 += vq[q].x - vq[q + 1].x;
        t2 += vq[q].y - vq[q + 1].y;
      }
      bre = t1;
      bim = t2;
    } else {
      const int pr = xr < N / 2 ? xr : N - xr;
      float s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f;
#pragma unroll
      for (int q = 0; q < KX; ++q) {
        const float2 e = Ex[(long)q * N + pr];
        s1 += vq[q].x * e.x;
        s2 += vq[q].y * e.y;
        s3 += vq[q].x * e.y;
        s4 += vq[q].y * e.x;
      }
      bre = xr == pr ? s1 - s2 : s1 + s2;
      bim = xr == pr ? s3 + s4 : s4 - s3;
    }
    lb[o][2 * ky] = (bf16)bre;
    lb[o][2 * ky + 1] = (bf16)bim;
    load_weights();
  }
  __syncthreads();
  bf16x8 ar[2];
#pragma unroll
  for (int to = 0; to < 2; ++to) ar[to] = *reinterpret_cast<const bf16x8*>(&lb[16 * to + i][8 * g]);
  // LIFT with at most four tiles per wave (N <= 512): every wave takes its
  // cells from the lifted row first, then the basis replaces the row before
  // the tile loop, its load latency hidden by the other waves' GEMMs
  // (block-uniform branch)
  const bool early = LIFT && ntile <= 32;
  if (early) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int t = w + 8 * u;
      t = t < ntile ? t : ntile - 1;
      const u32x2* sp = reinterpret_cast<const u32x2*>(lrow + (16 * t + i) * S + 8 * g);
      const u32x2 lo = sp[0], hi = sp[1];
      bpt[u] = __builtin_bit_cast(bf16x8, u32x4{lo[0], lo[1], hi[0], hi[1]});
    }
    __syncthreads();
    basis_to_lds<512>(Xt, lbas, N);
  }
  for (int k0 = 0; w + 8 * k0 < ntile; k0 += 4) {
    if (k0 > 0) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        int t = w + 8 * (k0 + u);
        t = t < ntile ? t : ntile - 1;
        const int y = 16 * t + i;
        if (!LIFT) bpt[u] = ld8(P + (bx * N + y) * C + 8 * g);
        bqt[u] = ld8(Q + (long)y * C + 8 * g);
      }
    }
    if (LIFT && !early) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = w + 8 * (k0 + u);
        if (t < ntile) {
          const u32x2* sp = reinterpret_cast<const u32x2*>(lrow + (16 * t + i) * S + 8 * g);
          const u32x2 lo = sp[0], hi = sp[1];
          bpt[u] = __builtin_bit_cast(bf16x8, u32x4{lo[0], lo[1], hi[0], hi[1]});
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = w + 8 * (k0 + u);
      if (t >= ntile) break;
      const int y = 16 * t + i;
#pragma unroll
      for (int to = 0; to < 2; ++to) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mma(aw[to], bpt[u], acc);
        acc = mma(ar[to], bqt[u], acc);
        const int oc0 = 16 * to + 4 * g;
        float r[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) r[v] = acc[v] + bs[to][v];
        if (pre_out) st4(pre_out + (bx * N + y) * C + oc0, r[0], r[1], r[2], r[3]);
#pragma unroll
        for (int v = 0; v < 4; ++v) r[v] = gelu(r[v]);
        st4(out + (bx * N + y) * C + oc0, r[0], r[1], r[2], r[3]);
#pragma unroll
        for (int v = 0; v < 4; ++v) lt[(oc0 + v) * basis_ld(N) + y] = (bf16)r[v];
      }
    }
  }
  __syncthreads();
  if (LIFT && !early) {  // the lifted row is dead: the basis takes its place
    basis_to_lds<512>(Xt, lbas, N);
    __syncthreads();
  }
  row_project<8>(lbas, basis_ld(N), lt, N, Anext + bx * C * C, red);
}

// k_cell_row<false, false> (B_x from R, the cells from P) with RPW
// consecutive rows per work-group, for N <= 512 (a wave owns at most four
// 16-cell tiles, the same ones in every row).  The basis F^T, the layer's
// weights and the F_y operands are staged once for the RPW rows, and the
// next row's cells and B_x are loaded before the current row's projection,
// so that the HBM reads of one row overlap the LDS work of the other.  The
// arithmetic of a row is k_cell_row's, operation for operation.
template <int RPW>
__global__ void __launch_bounds__(512) k_cell_rows(const bf16* __restrict__ P, const bf16* __restrict__ Q,
                                                   const bf16* __restrict__ A, const bf16* __restrict__ R,
                                                   const float* __restrict__ bias, bf16* __restrict__ pre_out,
                                                   bf16* __restrict__ out, const bf16* __restrict__ Xt,
                                                   float* __restrict__ Anext, int N, const int* done) {
  if (skip(done)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  bf16* lbas = reinterpret_cast<bf16*>(smem_raw);  // the basis F^T
  bf16* lt = lbas + C * basis_ld(N);               // h' of the row, channel-major
  __shared__ float red[4 * 64 * 4];
  __shared__ __align__(16) bf16 lb[C][C + 8];  // B_x as [o][k]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int ntile = N / 16;
  const long bx0 = (long)blockIdx.x * RPW;  // sample * N + x of the first row
  bf16x8 aw[2];
  float bs[2][4];
#pragma unroll
  for (int to = 0; to < 2; ++to) {
    aw[to] = ld8(A + (16 * to + i) * C + 8 * g);
#pragma unroll
    for (int v = 0; v < 4; ++v) bs[to][v] = bias[16 * to + 4 * g + v];
  }
  bf16x8 bpt[4], bqt[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    int t = w + 8 * u;
    t = t < ntile ? t : ntile - 1;
    const int y = 16 * t + i;
    bpt[u] = ld8(P + (bx0 * N + y) * C + 8 * g);
    bqt[u] = ld8(Q + (long)y * C + 8 * g);
  }
  const int bxo = (((tid >> 2) & 31)) * C + (tid & 3) * 8;  // this thread's 8 BF16 of B_x
  u32x4 bxv = *reinterpret_cast<const u32x4*>(R + bx0 * C * C + bxo);
  basis_to_lds<512>(Xt, lbas, N);
  // unrolled: as a loop the row-invariant addresses stay live across the
  // iterations (about 155 VGPRs, one work-group per CU); unrolled it needs
  // 109 at RPW 2 and 121 at RPW 4 and 8, two work-groups per CU
#pragma unroll
  for (int rr = 0; rr < RPW; ++rr) {
    const long bx = bx0 + rr;
    if (tid < C * C / 8) *reinterpret_cast<u32x4*>(&lb[tid >> 2][(tid & 3) * 8]) = bxv;
    __syncthreads();
    bf16x8 ar[2];
#pragma unroll
    for (int to = 0; to < 2; ++to) ar[to] = *reinterpret_cast<const bf16x8*>(&lb[16 * to + i][8 * g]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = w + 8 * u;
      if (t >= ntile) break;
      const int y = 16 * t + i;
#pragma unroll
      for (int to = 0; to < 2; ++to) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mma(aw[to], bpt[u], acc);
        acc = mma(ar[to], bqt[u], acc);
        const int oc0 = 16 * to + 4 * g;
        float r[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) r[v] = acc[v] + bs[to][v];
        if (pre_out) st4(pre_out + (bx * N + y) * C + oc0, r[0], r[1], r[2], r[3]);
#pragma unroll
        for (int v = 0; v < 4; ++v) r[v] = gelu(r[v]);
        st4(out + (bx * N + y) * C + oc0, r[0], r[1], r[2], r[3]);
#pragma unroll
        for (int v = 0; v < 4; ++v) lt[(oc0 + v) * basis_ld(N) + y] = (bf16)r[v];
      }
    }
    __syncthreads();
    if (rr + 1 < RPW) {  // the next row's operands (block-uniform branch)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        int t = w + 8 * u;
        t = t < ntile ? t : ntile - 1;
        bpt[u] = ld8(P + ((bx + 1) * N + 16 * t + i) * C + 8 * g);
      }
      bxv = *reinterpret_cast<const u32x4*>(R + (bx + 1) * C * C + bxo);
    }
    // waves of the upper K half leave row_project after its barrier; every
    // wave has read lt and red of this row before the next row's first
    // barrier, after which they are written again
    row_project<8>(lbas, basis_ld(N), lt, N, Anext + bx * C * C, red);
  }
}

// k_cell_row<false, true> (layer 0 of a batch) with RPW grid rows of one
// sample per block, N % RPW == 0 and N <= 512.  One LDS region holds the
// basis for all rows; the other holds a row lifted from (chi, v) as [y][S]
// until every wave has taken its cells, then h' channel-major for the row
// projection.  The weights, the F_y operands and the basis are loaded once;
// each next row's chi, v and B_x are loaded before the current row's
// projection.  The arithmetic of a row is k_cell_row's.
template <int RPW>
__global__ void __launch_bounds__(512) k_cell_rows_lift(LiftIn L, const bf16* __restrict__ Q,
                                                         const bf16* __restrict__ A, const bf16* __restrict__ R,
                                                         const float* __restrict__ bias, bf16* __restrict__ pre_out,
                                                         bf16* __restrict__ out, const bf16* __restrict__ Xt,
                                                         float* __restrict__ Anext, int N, const int* done) {
  if (skip(done)) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  bf16* lbas = reinterpret_cast<bf16*>(smem_raw);  // the basis F^T
  bf16* lx = lbas + C * basis_ld(N);               // the lifted row, then h'
  __shared__ float red[4 * 64 * 4];
  __shared__ __align__(16) bf16 lb[C][C + 8];  // B_x as [o][k]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i = lane & 15, g = lane >> 4;
  const int ntile = N / 16;
  const long bx0 = (long)blockIdx.x * RPW;
  const long ks = bx0 / N;  // sample of all RPW rows
  const int xr0 = (int)(bx0 - ks * N);
  bf16x8 aw[2];
  float bs[2][4];
#pragma unroll
  for (int to = 0; to < 2; ++to) {
    aw[to] = ld8(A + (16 * to + i) * C + 8 * g);
#pragma unroll
    for (int v = 0; v < 4; ++v) bs[to][v] = bias[16 * to + 4 * g + v];
  }
  bf16x8 bqt[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    int t = w + 8 * u;
    t = t < ntile ? t : ntile - 1;
    bqt[u] = ld8(Q + (long)(16 * t + i) * C + 8 * g);
  }
  const int bxo = ((tid >> 2) & 31) * C + (tid & 3) * 8;  // this thread's 8 BF16 of B_x
  u32x4 bxv = *reinterpret_cast<const u32x4*>(R + bx0 * C * C + bxo);
  LiftRegs lg;
  lift_load<512>(L, ks, xr0, N, tid, lg);
  basis_to_lds<512>(Xt, lbas, N);
#pragma unroll
  for (int rr = 0; rr < RPW; ++rr) {
    const long bx = bx0 + rr;
    if (rr) __syncthreads();  // the previous row's projection has read lx
    {
      float wl[32];
      lift_weights(L, wl);
      lift_store<512, false>(L, lg, wl, ks, xr0 + rr, N, tid, lx, nullptr, nullptr);
    }
    if (tid < C * C / 8) *reinterpret_cast<u32x4*>(&lb[tid >> 2][(tid & 3) * 8]) = bxv;
    __syncthreads();
    bf16x8 ar[2], bpt[4];
#pragma unroll
    for (int to = 0; to < 2; ++to) ar[to] = *reinterpret_cast<const bf16x8*>(&lb[16 * to + i][8 * g]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int t = w + 8 * u;
      t = t < ntile ? t : ntile - 1;
      const u32x2* sp = reinterpret_cast<const u32x2*>(lx + (16 * t + i) * S + 8 * g);
      const u32x2 lo = sp[0], hi = sp[1];
      bpt[u] = __builtin_bit_cast(bf16x8, u32x4{lo[0], lo[1], hi[0], hi[1]});
    }
    __syncthreads();  // every wave has its cells: lx takes h'
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = w + 8 * u;
      if (t >= ntile) break;
      const int y = 16 * t + i;
#pragma unroll
      for (int to = 0; to < 2; ++to) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mma(aw[to], bpt[u], acc);
        acc = mma(ar[to], bqt[u], acc);
        const int oc0 = 16 * to + 4 * g;
        float r[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) r[v] = acc[v] + bs[to][v];
        if (pre_out) st4(pre_out + (bx * N + y) * C + oc0, r[0], r[1], r[2], r[3]);
#pragma unroll
        for (int v = 0; v < 4; ++v) r[v] = gelu(r[v]);
        st4(out + (bx * N + y) * C + oc0, r[0], r[1], r[2], r[3]);
#pragma unroll
        for (int v = 0; v < 4; ++v) lx[(oc0 + v) * basis_ld(N) + y] = (bf16)r[v];
      }
    }
    __syncthreads();
    if (rr + 1 < RPW) {  // the next row's operands
      lift_load<512>(L, ks, xr0 + rr + 1, N, tid, lg);
      bxv = *reinterpret_cast<const u32x4*>(R + (bx + 1) * C * C + bxo);
    }
    row_project<8>(lbas, basis_ld(N), lx, N, Anext + bx * C * C, red);
  }
}

// Layer 3 (no activation) fused with the projection: per cell
//   h4 = [h3, F_y] [W; B_x] + b,  z = s_id vv / *vnrm + (h4 W_out + b_out) / s
// in FP32 (vnrm null: no scaling); h4 goes to hout as BF16 when given
// (training).  Sample k reads vv and writes z at k * sV, k * sZ, its norm at
// vnrm + k * sNrm bytes.  Grid as k_pixel.
__global__ void __launch_bounds__(256) k_cell_proj(const bf16* __restrict__ P, const bf16* __restrict__ Q,
                                                   const bf16* __restrict__ A, const bf16* __restrict__ R,
                                                   const float* __restrict__ bias, bf16* __restrict__ hout,
                                                   const float* __restrict__ Pp, const float2* __restrict__ vv,
                                                   long sV, const double* vnrm, long sNrm, float s,
                                                   float2* __restrict__ z, long sZ, int N, const int* done) {
  if (skip(done)) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const long wv = (long)blockIdx.x * 4 + w;  // every block is full: waves % 4 == 0
  const int segs = N / 64;
  const long bx = wv / segs;
  const int y0 = (int)(wv % segs) * 64;
  const int i = lane & 15, g = lane >> 4;
  const long ks = bx / N;                 // sample
  const long rin = (bx - ks * N) * N;     // the row's first cell in the sample
  vv += ks * sV;
  z += ks * sZ;
  // every load of the wave is issued before the first use of any: the cell
  // operands, the vector's values, the weights, then the vector's norm (read
  // through a pointer that is never null, so that no branch waits for the
  // loads before it)
  bf16x8 bpt[4], bqt[4];
  float2 vx[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int y = y0 + 16 * t + i;
    bpt[t] = ld8(P + (bx * N + y) * C + 8 * g);
    bqt[t] = ld8(Q + (long)y * C + 8 * g);
    vx[t] = vv[rin + y];
  }
  bf16x8 aw[2], ar[2];
  float bs[2][4], wo[2][4][2];
  const float* Wout = Pp + kOffOut;
#pragma unroll
  for (int to = 0; to < 2; ++to) {
    aw[to] = ld8(A + (16 * to + i) * C + 8 * g);
    ar[to] = ld8(R + (bx * C + 16 * to + i) * C + 8 * g);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int o = 16 * to + 4 * g + v;
      bs[to][v] = bias[o];
      wo[to][v][0] = Wout[2 * o];
      wo[to][v][1] = Wout[2 * o + 1];
    }
  }
  const double d = *(vnrm ? boff(vnrm, ks * sNrm) : &kOne);  // no norm: 1
  const float sid = Pp[kOffSkip] * (d > 0.0 ? (float)(1.0 / d) : 0.f);
  const float b0 = Wout[2 * C], b1 = Wout[2 * C + 1], is = 1.f / s;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const long p = bx * N + y0 + 16 * t + i;
    float o0 = 0.f, o1 = 0.f;
#pragma unroll
    for (int to = 0; to < 2; ++to) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
      acc = mma(aw[to], bpt[t], acc);
      acc = mma(ar[to], bqt[t], acc);
      float r[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        r[v] = acc[v] + bs[to][v];
        o0 += r[v] * wo[to][v][0];
        o1 += r[v] * wo[to][v][1];
      }
      if (hout) st4(hout + p * C + 16 * to + 4 * g, r[0], r[1], r[2], r[3]);
    }
    // the four lane groups g hold channels 4g..4g+3 (+16) of cell i
    o0 += __shfl_xor(o0, 16, 64);
    o1 += __shfl_xor(o1, 16, 64);
    o0 += __shfl_xor(o0, 32, 64);
    o1 += __shfl_xor(o1, 32, 64);
    if (g == 0) {
      const float2 x = vx[t];
      z[rin + y0 + 16 * t + i] = make_float2(sid * x.x + (o0 + b0) * is, sid * x.y + (o1 + b1) * is);
    }
  }
}


// Block sums of NV per-thread values (every thread of the 256 passes vals);
// thread k < NV writes the block's sum of value k to dst[k].
template <int NV>
__device__ inline void block_sums(const float* vals, float* dst) {
  __shared__ float red[4][NV];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const float t = wave_sum_f(vals[k]);
    if (lane == 0) red[w][k] = t;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NV; k += 256) dst[k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
}

// projection backward: dz = dL/dz (re, im packed) -> d = dz / s;
//   dh[p][c] = d0 W_out[c][0] + d1 W_out[c][1]   (BF16)
//   part[blk] = (sum_p h[p][c] d_j  [c][j],  sum_p d_j,  sum_p Re(conj(dz) vv))
__global__ void __launch_bounds__(256) k_proj_bwd(const float2* __restrict__ dz, const bf16* __restrict__ h,
                                                  const float* __restrict__ P, const float2* __restrict__ v,
                                                  float s, bf16* __restrict__ dh, float* __restrict__ part,
                                                  long total) {
  const float* W = P + kOffOut;
  float acc[2 * C + 3];
#pragma unroll
  for (int k = 0; k < 2 * C + 3; ++k) acc[k] = 0.f;
  const float is = 1.f / s;
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < total; p += (long)gridDim.x * 256) {
    const float2 d = dz[p];
    const float d0 = d.x * is, d1 = d.y * is;
#pragma unroll
    for (int c8 = 0; c8 < C; c8 += 8) {
      const bf16x8 hv = ld8(h + p * C + c8);
      float r[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int c = c8 + k;
        const float x = (float)hv[k];
        acc[2 * c] += x * d0;
        acc[2 * c + 1] += x * d1;
        r[k] = d0 * W[2 * c] + d1 * W[2 * c + 1];
      }
      st4(dh + p * C + c8, r[0], r[1], r[2], r[3]);
      st4(dh + p * C + c8 + 4, r[4], r[5], r[6], r[7]);
    }
    acc[2 * C] += d0;
    acc[2 * C + 1] += d1;
    const float2 vv = v[p];
    acc[2 * C + 2] += d.x * vv.x + d.y * vv.y;
  }
  block_sums<2 * C + 3>(acc, part + (long)blockIdx.x * (2 * C + 3));
}

// lift backward: part[blk] = (sum_p a_i dh0[p][c]  [i][c],  sum_p dh0[p][c])
__global__ void __launch_bounds__(256) k_lift_bwd(const bf16* __restrict__ dh0, const float* __restrict__ chi,
                                                  float chi_scale, const float2* __restrict__ v, float s,
                                                  float* __restrict__ part, long total) {
  float acc[4 * C];
#pragma unroll
  for (int k = 0; k < 4 * C; ++k) acc[k] = 0.f;
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < total; p += (long)gridDim.x * 256) {
    const float2 vv = v[p];
    const float a[3] = {chi[p] * chi_scale, vv.x * s, vv.y * s};
#pragma unroll
    for (int c8 = 0; c8 < C; c8 += 8) {
      const bf16x8 hv = ld8(dh0 + p * C + c8);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const float d = (float)hv[k];
        acc[c8 + k] += a[0] * d;
        acc[C + c8 + k] += a[1] * d;
        acc[2 * C + c8 + k] += a[2] * d;
        acc[3 * C + c8 + k] += d;
      }
    }
  }
  block_sums<4 * C>(acc, part + (long)blockIdx.x * 4 * C);
}

// out[k] = sum_g part[g * width + k]: 64 outputs per block, 16 strided
// slices of g per output combined in a fixed order.
__global__ void __launch_bounds__(1024) k_reduce_part(const float* __restrict__ part, int G, int width,
                                                      float* __restrict__ out) {
  __shared__ float red[16][65];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + tx;
  float s = 0.f;
  if (k < width)
    for (int gi = ty; gi < G; gi += 16) s += part[(long)gi * width + k];
  red[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && k < width) {
    float t = 0.f;
    for (int q = 0; q < 16; ++q) t += red[q][tx];
    out[k] = t;
  }
}

// GEMM weights to BF16: Wt[l][o][c] = W_l[c][o] (forward A operand),
// Wn[l][c][o] = W_l[c][o] (backward A operand).
__global__ void __launch_bounds__(256) k_prep(const float* __restrict__ P, bf16* __restrict__ Wt,
                                              bf16* __restrict__ Wn) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= 4 * C * C) return;
  const int l = e / (C * C), r = e % (C * C), c = r / C, o = r % C;
  const float w = P[layer_off(l) + r];
  Wn[e] = (bf16)w;
  Wt[l * C * C + o * C + c] = (bf16)w;
}

}  // namespace fno
}  // namespace cgm

using namespace cgm;
using namespace cgm::fno;

struct cgm_fno {
  cgm_context* ctx = nullptr;
  int device = 0;  // kept here so the network can be destroyed after its context
  int N = 0, B = 0;
  long n = 0;
  float chi_scale = 1.f;
  const float* params = nullptr;  // caller's FP32 parameters (device)
  bf16* Wt = nullptr;             // 4 x 32 x 32
  bf16* Wn = nullptr;
  bf16* Ft = nullptr;             // 32 x N
  bf16* Fy = nullptr;             // N x 32
  float2* Ex = nullptr;           // [32 kxi][N]: e^{2 pi i kx x/N}
  bf16* h[5] = {};                // activations [B][n][32]
  bf16* pre[3] = {};              // pre-activations (training)
  bf16* dh[2] = {};               // backward ping-pong
  float* A = nullptr;             // [B][N][32][32]
  float2* upart = nullptr;        // [B][N/XC][32][16][32]
  float2* U[4] = {};              // [B][32][16][32] (training)
  float2* dV = nullptr;           // [B][32][16][32]
  float2* Vm = nullptr;           // [B][32][16][32] mixed modes
  bf16* Bt = nullptr;             // [B][N][32][32]
  float* part = nullptr;          // reduction partials
  int Gp = 0;                     // blocks of the reducing kernels
  // saved by the last training forward
  int fwd_batch = 0;
  const float* chi_in = nullptr;
  const float2* v_in = nullptr;
};

namespace {

int fno_free(cgm_fno* f) {
  void* ptrs[] = {f->Wt, f->Wn, f->Ft, f->Fy, f->Ex, f->h[0], f->h[1], f->h[2], f->h[3], f->h[4],
                  f->pre[0], f->pre[1], f->pre[2], f->dh[0], f->dh[1], f->A, f->upart, f->U[0],
                  f->U[1], f->U[2], f->U[3], f->dV, f->Vm, f->Bt, f->part};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  return CGM_SUCCESS;
}

int rowproj_grid(int rows) { return rows < 256 ? rows : 256; }

template <bool WZ>
int launch_rowproj(cgm_fno* f, hipStream_t st, const bf16* Y, const bf16* Z, float* out, int B,
                   const int* done) {
  const int rows = B * f->N;
  const int G = WZ ? f->Gp : rows;
  const size_t lds = (size_t)(WZ ? 2 : 1) * f->N * S * sizeof(bf16);
  hipLaunchKernelGGL((k_rowproj<WZ>), dim3(G), dim3(256), lds, st, f->Ft, Y, Z, out, f->part, f->N, rows,
                     done);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

template <int EPI>
int launch_pixel(cgm_fno* f, hipStream_t st, const bf16* P, const bf16* A, const bf16* R, const float* bias,
                 bf16* pre_out, const bf16* aux, bf16* out, int B, const int* done) {
  const long waves = (long)B * f->n / 64;  // N % 64 == 0, so waves % 4 == 0
  CGM_CHECK(waves % 4 == 0, CGM_ERROR_UNSUPPORTED, "cell GEMM needs N %% 64 == 0");
  hipLaunchKernelGGL((k_pixel<EPI>), dim3((unsigned)(waves / 4)), dim3(256), 0, st, P, f->Fy, A, R, bias, pre_out,
                     aux, out, f->N, done);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

void launch_invx(cgm_fno* f, hipStream_t st, int B, const int* done) {
  if (B >= 8)  // every work-group reads its vector's 128 KiB of modes: fewer, larger ones
    hipLaunchKernelGGL(k_invx<8>, dim3(f->N / 16, 1, B), dim3(512), 0, st, f->Vm, f->Bt, f->Ex, f->N, done);
  else if (B >= 4)
    hipLaunchKernelGGL(k_invx<4>, dim3(f->N / 8, 1, B), dim3(512), 0, st, f->Vm, f->Bt, f->Ex, f->N, done);
  else
    hipLaunchKernelGGL(k_invx<1>, dim3(f->N / 2, 1, B), dim3(512), 0, st, f->Vm, f->Bt, f->Ex, f->N, done);
}

// x-DFT of A (B vectors) into partials over chunks of 64 rows, whatever B
// (the same partial sums for one vector and for a batch); returns the number
// of chunks
int launch_dftx(cgm_fno* f, hipStream_t st, int scaled, int B, const int* done) {
  hipLaunchKernelGGL(k_dftx<64>, dim3(f->N / 64, KY, B), dim3(256), 0, st, f->A, f->upart, f->Ex, f->N, scaled, done);
  return f->N / 64;
}

// spectral path of one layer: A (row projections) -> Bt (per-row mixing matrices)
int spectral(cgm_fno* f, hipStream_t st, const float2* W, float2* usave, bool bwd, int B, const int* done,
             bool invx = true) {
  const int nch = launch_dftx(f, st, bwd ? 1 : 0, B, done);
  if (!bwd)
    hipLaunchKernelGGL((k_mix<false>), dim3(KY, KX / KXG, B), dim3(256), 0, st, f->upart, nch, W, usave, f->Vm,
                       f->N, done);
  else
    hipLaunchKernelGGL((k_mix<true>), dim3(KY, KX / KXG, B), dim3(256), 0, st, f->upart, nch, W, usave, f->Vm,
                       f->N, done);
  if (invx)
    launch_invx(f, st, B, done);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

}  // namespace

namespace cgm {

// Two grid rows per block where N allows it (N even, N <= 512) and the
// grid still holds at least 1024 blocks.
inline bool two_rows(int N, int rows) { return N % 2 == 0 && N <= 512 && rows / 2 >= 1024; }

// The lift with layer 0's first row projection (k_lift_row / k_lift_rows):
// four rows per block where the grid still holds 1024 blocks, else two.
void launch_lift(const cgm_fno* f, hipStream_t st, int rows, const LiftIn& L, float2* vw, bf16* h0,
                 const int* done) {
  const size_t lds = row_kernel_lds(f->N);
  if (f->N % 4 == 0 && f->N <= 512 && rows / 4 >= 1024)
    hipLaunchKernelGGL(k_lift_rows<4>, dim3(rows / 4), dim3(512), lds, st, L, vw, h0, f->Ft, f->A, f->N, done);
  else if (two_rows(f->N, rows))
    hipLaunchKernelGGL(k_lift_rows<2>, dim3(rows / 2), dim3(512), lds, st, L, vw, h0, f->Ft, f->A, f->N, done);
  else
    hipLaunchKernelGGL(k_lift_row, dim3(rows), dim3(512), lds, st, L, vw, h0, f->Ft, f->A, f->N, done);
}

// Layer 0's cell GEMM on B > 1 samples, lifting its rows from (chi, v).
void launch_cell_rows_lift(const cgm_fno* f, hipStream_t st, int rows, const LiftIn& L, const bf16* W,
                           const float* bias, bf16* pre, bf16* out, const int* done) {
  const size_t lds = row_kernel_lds(f->N);
  if (f->N % 4 == 0 && f->N <= 512 && rows / 4 >= 1024)
    hipLaunchKernelGGL(k_cell_rows_lift<4>, dim3(rows / 4), dim3(512), lds, st, L, f->Fy, W, f->Bt, bias, pre, out,
                       f->Ft, f->A, f->N, done);
  else if (two_rows(f->N, rows))
    hipLaunchKernelGGL(k_cell_rows_lift<2>, dim3(rows / 2), dim3(512), lds, st, L, f->Fy, W, f->Bt, bias, pre, out,
                       f->Ft, f->A, f->N, done);
  else
    hipLaunchKernelGGL((k_cell_row<false, true>), dim3(rows), dim3(512), lds, st, nullptr, L, f->Fy, W, f->Bt,
                       nullptr, nullptr, bias, pre, out, f->Ft, f->A, f->N, done);
}

// Cell GEMM of layers 1-2 on B > 1 samples (B_x from R): the most rows per
// work-group (8, 4 or 2) that leave at least 1024 work-groups.
void launch_cell_rows(const cgm_fno* f, hipStream_t st, int rows, const bf16* P, const bf16* W, const float* bias,
                      bf16* pre, bf16* out, const int* done) {
  const size_t lds = row_kernel_lds(f->N);
  if (f->N % 8 == 0 && f->N <= 512 && rows / 8 >= 1024)
    hipLaunchKernelGGL(k_cell_rows<8>, dim3(rows / 8), dim3(512), lds, st, P, f->Fy, W, f->Bt, bias, pre, out, f->Ft,
                       f->A, f->N, done);
  else if (f->N % 4 == 0 && f->N <= 512 && rows / 4 >= 1024)
    hipLaunchKernelGGL(k_cell_rows<4>, dim3(rows / 4), dim3(512), lds, st, P, f->Fy, W, f->Bt, bias, pre, out, f->Ft,
                       f->A, f->N, done);
  else if (two_rows(f->N, rows))
    hipLaunchKernelGGL(k_cell_rows<2>, dim3(rows / 2), dim3(512), lds, st, P, f->Fy, W, f->Bt, bias, pre, out, f->Ft,
                       f->A, f->N, done);
  else
    hipLaunchKernelGGL((k_cell_row<false, false>), dim3(rows), dim3(512), lds, st, P, LiftIn{}, f->Fy, W, f->Bt,
                       nullptr, nullptr, bias, pre, out, f->Ft, f->A, f->N, done);
}

// Forward pass on B unit-norm vectors (or one lazily normalised vector with
// nrm/vw).  train keeps activations for fno_backward.
int fno_forward_io(cgm_fno* f, hipStream_t st, const FnoIO& io, int B, bool train, const int* done) {
  const float s = (float)f->N;
  const int rows = B * f->N;
  const size_t lds = row_kernel_lds(f->N);
  const bool fold = B == 1;
  int rc;
  LiftIn L{io.chi, io.sChi, f->chi_scale, io.v, io.sV, io.nrm, io.sNrm, s, f->params};
  launch_lift(f, st, rows, L, io.vw, train ? f->h[0] : nullptr, done);
  CGM_HIP(hipGetLastError());
  const float2* vv = io.vw ? io.vw : io.v;
  // layer 0 lifts its rows again from v' (normalised by k_lift_row into vw
  // when given) instead of reading h0
  LiftIn L0 = L;
  if (io.vw) {
    L0.v = io.vw;
    L0.nrm = nullptr;
  }
  for (int l = 0; l < 4; ++l) {
    const float* Lp = f->params + layer_off(l);
    // x-DFT and mixing; for one sample the inverse x-DFT of each row of
    // layers 0-2 runs in its cell GEMM, otherwise k_invx forms every B_x once
    const bool fold_l = fold && l < 3;
    if ((rc = spectral(f, st, (const float2*)(Lp + C * C + C), train ? f->U[l] : nullptr, false, B, done,
                       !fold_l)))
      return rc;
    const bf16* Wl = f->Wt + l * C * C;
    if (l > 0 && l < 3 && !fold) {
      launch_cell_rows(f, st, rows, f->h[l], Wl, Lp + C * C, train ? f->pre[l] : nullptr, f->h[l + 1], done);
    } else if (l == 0 && !fold) {
      launch_cell_rows_lift(f, st, rows, L0, Wl, Lp + C * C, train ? f->pre[0] : nullptr, f->h[1], done);
    } else if (l < 3) {
      auto kern = l == 0 ? k_cell_row<true, true> : k_cell_row<true, false>;
      hipLaunchKernelGGL(kern, dim3(rows), dim3(512), lds, st, f->h[l], L0, f->Fy, Wl, f->Bt, f->Vm, f->Ex,
                         Lp + C * C, train ? f->pre[l] : nullptr, f->h[l + 1], f->Ft, f->A, f->N, done);
    } else {
      const long waves = (long)B * f->n / 64;
      CGM_CHECK(waves % 4 == 0, CGM_ERROR_UNSUPPORTED, "cell GEMM needs N %% 64 == 0");
      hipLaunchKernelGGL(k_cell_proj, dim3((unsigned)(waves / 4)), dim3(256), 0, st, f->h[3], f->Fy, Wl, f->Bt,
                         Lp + C * C, train ? f->h[4] : nullptr, f->params, vv, io.sV, io.vw ? nullptr : io.nrm,
                         io.sNrm, s, io.z, io.sZ, f->N, done);
    }
    CGM_HIP(hipGetLastError());
  }
  if (train) {
    f->fwd_batch = B;
    f->chi_in = io.chi;
    f->v_in = vv;
  }
  return CGM_SUCCESS;
}

int fno_forward(cgm_fno* f, hipStream_t st, const float* chi, const float2* v, const double* nrm, float2* vw,
                float2* z, int B, bool train, const int* done) {
  FnoIO io;
  io.chi = chi;
  io.sChi = f->n;
  io.v = v;
  io.sV = f->n;
  io.nrm = nrm;
  io.sNrm = 0;
  io.vw = vw;
  io.z = z;
  io.sZ = f->n;
  return fno_forward_io(f, st, io, B, train, done);
}

int fno_batch(const cgm_fno* f) { return f->B; }

cgm_context* fno_context(const cgm_fno* f) { return f->ctx; }

}  // namespace cgm

static int fno_backward(cgm_fno* f, const float2* dz, float* grad) {
  hipStream_t st = f->ctx->stream;
  const int B = f->fwd_batch;
  const long total = (long)B * f->n;
  const float s = (float)f->N;
  int rc;
  const int Gr = 256;
  // projection
  hipLaunchKernelGGL(k_proj_bwd, dim3(Gr), dim3(256), 0, st, dz, f->h[4], f->params, f->v_in, s, f->dh[0], f->part,
                     total);
  hipLaunchKernelGGL(k_reduce_part, dim3(2), dim3(1024), 0, st, f->part, Gr, 2 * C + 3, grad + kOffOut);
  CGM_HIP(hipGetLastError());
  // layer 3 has no activation: dPre_3 = dh4
  int cur = 0;
  for (int l = 3; l >= 0; --l) {
    const float* Lp = f->params + layer_off(l);
    float* Lg = grad + layer_off(l);
    const bf16* dpre = f->dh[cur];
    // g_B rows (basis F), dW_l and db_l partials
    if ((rc = launch_rowproj<true>(f, st, dpre, f->h[l], f->A, B, nullptr))) return rc;
    hipLaunchKernelGGL(k_reduce_part, dim3((C * C + C + 63) / 64), dim3(1024), 0, st, f->part, f->Gp, C * C + C,
                       Lg);
    // spectral adjoint: dV (saved), dA rows into Bt
    if ((rc = spectral(f, st, (const float2*)(Lp + C * C + C), f->dV, true, B, nullptr))) return rc;
    hipLaunchKernelGGL(k_specgrad, dim3(KX * KY * C * C / 256), dim3(256), 0, st, f->U[l], f->dV,
                       (float2*)(Lg + C * C + C), B);
    CGM_HIP(hipGetLastError());
    const bf16* Wl = f->Wn + l * C * C;
    if (l > 0)
      rc = launch_pixel<2>(f, st, dpre, Wl, f->Bt, nullptr, nullptr, f->pre[l - 1], f->dh[cur ^ 1], B, nullptr);
    else
      rc = launch_pixel<0>(f, st, dpre, Wl, f->Bt, nullptr, nullptr, nullptr, f->dh[cur ^ 1], B, nullptr);
    if (rc) return rc;
    cur ^= 1;
  }
  hipLaunchKernelGGL(k_lift_bwd, dim3(Gr), dim3(256), 0, st, f->dh[cur], f->chi_in, f->chi_scale, f->v_in, s,
                     f->part, total);
  hipLaunchKernelGGL(k_reduce_part, dim3(2), dim3(1024), 0, st, f->part, Gr, 4 * C, grad);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

extern "C" int cgm_fno_num_params(size_t* n) {
  CGM_CHECK(n, CGM_ERROR_INVALID_VALUE, "null argument");
  *n = (size_t)kNumParams;
  return CGM_SUCCESS;
}

extern "C" int cgm_fno_create(cgm_ctx_t c, int batch, float chi_scale, cgm_fno_t* out) {
  CGM_CHECK(c && out, CGM_ERROR_INVALID_VALUE, "null argument");
  *out = nullptr;
  CGM_CHECK(batch >= 1, CGM_ERROR_INVALID_VALUE, "batch %d < 1", batch);
  const int N = c->N;
  CGM_CHECK(N >= 64 && N <= 1024 && N % 64 == 0, CGM_ERROR_UNSUPPORTED,
            "the FNO preconditioner needs 64 <= N <= 1024 with N %% 64 == 0 (N = %d)", N);
  CGM_HIP(hipSetDevice(c->device));
  cgm_fno* f = new cgm_fno();
  f->ctx = c;
  f->device = c->device;
  f->N = N;
  f->B = batch;
  f->n = (long)N * N;
  f->chi_scale = chi_scale;
  const size_t act = sizeof(bf16) * (size_t)batch * f->n * C;
  int rc = CGM_SUCCESS;
  auto alloc = [&](void** p, size_t bytes) {
    if (rc) return;
    if (hipMalloc(p, bytes) != hipSuccess) {
      cgm_set_last_error("cgm_fno_create: hipMalloc of %zu bytes failed", bytes);
      rc = CGM_ERROR_ALLOC;
    }
  };
  alloc((void**)&f->Wt, sizeof(bf16) * 4 * C * C);
  alloc((void**)&f->Wn, sizeof(bf16) * 4 * C * C);
  alloc((void**)&f->Ft, sizeof(bf16) * C * N);
  alloc((void**)&f->Fy, sizeof(bf16) * C * N);
  alloc((void**)&f->Ex, sizeof(float2) * KX * N);
  for (int l = 0; l < 5; ++l) alloc((void**)&f->h[l], act);
  for (int l = 0; l < 3; ++l) alloc((void**)&f->pre[l], act);
  for (int l = 0; l < 2; ++l) alloc((void**)&f->dh[l], act);
  alloc((void**)&f->A, sizeof(float) * (size_t)batch * N * C * C);
  alloc((void**)&f->upart, sizeof(float2) * (size_t)batch * (N / XC) * KX * KY * C);
  for (int l = 0; l < 4; ++l) alloc((void**)&f->U[l], sizeof(float2) * (size_t)batch * KX * KY * C);
  alloc((void**)&f->dV, sizeof(float2) * (size_t)batch * KX * KY * C);
  alloc((void**)&f->Vm, sizeof(float2) * (size_t)batch * KX * KY * C);
  alloc((void**)&f->Bt, sizeof(bf16) * (size_t)batch * N * C * C);
  f->Gp = rowproj_grid(batch * N);
  size_t pw = (size_t)f->Gp * (C * C + C);
  if (pw < (size_t)256 * 4 * C) pw = (size_t)256 * 4 * C;
  alloc((void**)&f->part, sizeof(float) * pw);
  if (rc) {
    fno_free(f);
    delete f;
    return rc;
  }
  // column basis and row twiddles (computed in double, stored rounded)
  std::vector<bf16> ft((size_t)C * N), fy((size_t)C * N);
  std::vector<float2> ex((size_t)KX * N);
  for (int y = 0; y < N; ++y) {
    for (int ky = 0; ky < KY; ++ky) {
      const double th = 2.0 * M_PI * (double)((long)ky * y % N) / N;
      const float cs = (float)cos(th), sn = (float)-sin(th);
      ft[(size_t)(2 * ky) * N + y] = (bf16)cs;
      ft[(size_t)(2 * ky + 1) * N + y] = (bf16)sn;
      fy[(size_t)y * C + 2 * ky] = (bf16)cs;
      fy[(size_t)y * C + 2 * ky + 1] = (bf16)sn;
    }
    for (int q = 0; q < KX; ++q) {
      const long kx = q < 16 ? q : N - KX + q;
      const double ph = 2.0 * M_PI * (double)(kx * y % N) / N;
      ex[(size_t)q * N + y] = make_float2((float)cos(ph), (float)sin(ph));
    }
  }
  CGM_HIP(hipMemcpy(f->Ft, ft.data(), sizeof(bf16) * ft.size(), hipMemcpyHostToDevice));
  CGM_HIP(hipMemcpy(f->Fy, fy.data(), sizeof(bf16) * fy.size(), hipMemcpyHostToDevice));
  CGM_HIP(hipMemcpy(f->Ex, ex.data(), sizeof(float2) * ex.size(), hipMemcpyHostToDevice));
  const size_t lds2 = (size_t)2 * N * S * sizeof(bf16);
  CGM_HIP(hipFuncSetAttribute((const void*)k_rowproj<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2));
  CGM_HIP(hipFuncSetAttribute((const void*)k_rowproj<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(lds2 / 2)));
  CGM_HIP(hipFuncSetAttribute((const void*)k_lift_row, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)row_kernel_lds(N)));
  const void* rowk[] = {(const void*)k_cell_row<true, false>, (const void*)k_cell_row<false, false>,
                        (const void*)k_cell_row<true, true>,  (const void*)k_cell_row<false, true>,
                        (const void*)k_cell_rows<2>,          (const void*)k_cell_rows<4>,
                        (const void*)k_cell_rows<8>,
                        (const void*)k_cell_rows_lift<2>,     (const void*)k_cell_rows_lift<4>,
                        (const void*)k_lift_rows<2>,          (const void*)k_lift_rows<4>};
  for (const void* k : rowk)
    CGM_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)row_kernel_lds(N)));
  *out = f;
  return CGM_SUCCESS;
}

extern "C" int cgm_fno_destroy(cgm_fno_t f) {
  if (!f) return CGM_SUCCESS;
  (void)hipSetDevice(f->device);
  fno_free(f);
  delete f;
  return CGM_SUCCESS;
}

extern "C" int cgm_fno_set_params(cgm_fno_t f, const float* params) {
  CGM_CHECK(f && params, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(((uintptr_t)params & 15) == 0, CGM_ERROR_INVALID_VALUE, "parameters must be 16-byte aligned");
  CGM_HIP(hipSetDevice(f->ctx->device));
  f->params = params;
  hipLaunchKernelGGL(k_prep, dim3(4 * C * C / 256), dim3(256), 0, f->ctx->stream, params, f->Wt, f->Wn);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

extern "C" int cgm_fno_forward(cgm_fno_t f, const float* chi, const cgm_complex64* v, cgm_complex64* z, int batch,
                               int train) {
  CGM_CHECK(f && chi && v && z, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(f->params, CGM_ERROR_INVALID_VALUE, "cgm_fno_set_params has not been called");
  CGM_CHECK(batch >= 1 && batch <= f->B, CGM_ERROR_INVALID_VALUE, "batch %d outside [1, %d]", batch, f->B);
  CGM_CHECK((const void*)v != (const void*)z, CGM_ERROR_INVALID_VALUE, "v and z must not alias");
  CGM_HIP(hipSetDevice(f->ctx->device));
  return fno_forward(f, f->ctx->stream, chi, (const float2*)v, nullptr, nullptr, (float2*)z, batch, train != 0,
                     nullptr);
}

extern "C" int cgm_fno_backward(cgm_fno_t f, const cgm_complex64* dz, float* grad) {
  CGM_CHECK(f && dz && grad, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(f->fwd_batch > 0, CGM_ERROR_INVALID_VALUE, "cgm_fno_backward needs a preceding training forward");
  CGM_HIP(hipSetDevice(f->ctx->device));
  return fno_backward(f, (const float2*)dz, grad);
}

// Average device time (ms, HIP events on the context stream, reps launches
// after two warm-ups) of each kernel of the network on `batch` vectors
// (contiguous [batch][N*N] chi, v, z): ms[0] lift with the first row projection (k_lift_row),
// [1] row projection alone (k_rowproj, used by the backward pass), [2] x-DFT,
// [3] mixing, [4] inverse x-DFT, [5] cell GEMM with GELU and the next row
// projection (k_cell_row of layers 1-2, forming B_x itself for batch 1), [6]
// last cell GEMM with the projection (k_cell_proj), [7] one whole forward
// pass, [8] the cell GEMM of layer 0 (k_cell_row lifting its row from chi
// and v).  The lift of [0] keeps h0 in LDS only, as in inference.
extern "C" int cgm_fno_profile(cgm_fno_t f, const float* chi, const cgm_complex64* v, cgm_complex64* z, int batch,
                               int reps, float* ms) {
  CGM_CHECK(f && chi && v && z && ms && reps > 0, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(batch >= 1 && batch <= f->B, CGM_ERROR_INVALID_VALUE, "batch %d outside [1, %d]", batch, f->B);
  CGM_CHECK(f->params, CGM_ERROR_INVALID_VALUE, "cgm_fno_set_params has not been called");
  CGM_HIP(hipSetDevice(f->ctx->device));
  hipStream_t st = f->ctx->stream;
  const int B = batch;
  const long total = (long)B * f->n;
  const unsigned rows = (unsigned)(B * f->N);
  const float s = (float)f->N;
  int nch = 0;
  const float* Lp = f->params + layer_off(0);
  const float2* W = (const float2*)(Lp + C * C + C);
  const size_t lds = row_kernel_lds(f->N);
  hipEvent_t e0, e1;
  CGM_HIP(hipEventCreate(&e0));
  CGM_HIP(hipEventCreate(&e1));
  const LiftIn L{chi, f->n, f->chi_scale, (const float2*)v, f->n, nullptr, 0L, s, f->params};
  int rc = CGM_SUCCESS;
  for (int k = 0; k < 9 && !rc; ++k) {
    for (int r = -2; r < reps && !rc; ++r) {
      if (r == 0) CGM_HIP(hipEventRecord(e0, st));
      switch (k) {
        case 0:  // the first warm-up also stores h0, which cases 1 and 5 read
          launch_lift(f, st, rows, L, nullptr, r == -2 ? f->h[0] : nullptr, nullptr);
          break;
        case 1: rc = launch_rowproj<false>(f, st, f->h[0], nullptr, f->A, B, nullptr); break;
        case 2: nch = launch_dftx(f, st, 0, B, nullptr); break;
        case 3:
          hipLaunchKernelGGL((k_mix<false>), dim3(KY, KX / KXG, B), dim3(256), 0, st, f->upart, nch, W, nullptr,
                             f->Vm, f->N, nullptr);
          break;
        case 4:
          launch_invx(f, st, B, nullptr);
          break;
        case 5:
          if (B == 1)
            hipLaunchKernelGGL((k_cell_row<true, false>), dim3(rows), dim3(512), lds, st, f->h[0], L, f->Fy, f->Wt,
                               f->Bt, f->Vm, f->Ex, Lp + C * C, nullptr, f->h[1], f->Ft, f->A, f->N, nullptr);
          else
            launch_cell_rows(f, st, rows, f->h[0], f->Wt, Lp + C * C, nullptr, f->h[1], nullptr);
          break;
        case 8:
          if (B == 1)
            hipLaunchKernelGGL((k_cell_row<true, true>), dim3(rows), dim3(512), lds, st, nullptr, L, f->Fy, f->Wt,
                               f->Bt, f->Vm, f->Ex, Lp + C * C, nullptr, f->h[1], f->Ft, f->A, f->N, nullptr);
          else
            launch_cell_rows_lift(f, st, rows, L, f->Wt, Lp + C * C, nullptr, f->h[1], nullptr);
          break;
        case 6:
          hipLaunchKernelGGL(k_cell_proj, dim3((unsigned)(total / 256)), dim3(256), 0, st, f->h[3], f->Fy, f->Wt,
                             f->Bt, Lp + C * C, nullptr, f->params, (const float2*)v, f->n, nullptr, 0L, s,
                             (float2*)z, f->n, f->N, nullptr);
          break;
        default:
          rc = fno_forward(f, st, chi, (const float2*)v, nullptr, nullptr, (float2*)z, B, false, nullptr);
          break;
      }
      if (hipGetLastError() != hipSuccess) rc = CGM_ERROR_DEVICE;
    }
    CGM_HIP(hipEventRecord(e1, st));
    CGM_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    CGM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[k] = t / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}
