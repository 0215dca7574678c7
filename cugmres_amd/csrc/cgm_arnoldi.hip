// Fused Arnoldi steps with lagged CGS2 (delayed reorthogonalisation) for
// grids N = 2^k (64 <= N <= 2048) and N = 3 * 2^k (96 <= N <= 768) with
// restart <= 64.
//
// CGS2 (two classical Gram-Schmidt passes per step) replaces the reference's
// modified Gram-Schmidt (B/CUDA_GMRES/GMRES.cuh:289-303).  Here the second
// pass of step j-1 is delayed and merged into the single global reduction of
// step j (the DCGS2 scheme of Swirydowicz, Langou, Ananthan, Yang and Thomas,
// "Low synchronization Gram-Schmidt and generalized minimal residual
// algorithms", Numer. Linear Algebra Appl. 28 (2021) e2343).  With u_j the
// once-orthogonalised, unnormalised vector of step j and Q_j = [q_0..q_{j-1}]:
//
//   z = A u_j                                          (FFT matvec)
//   s = Q_j^H u_j,  t = Q_j^H z,  alpha = u_j^H u_j,  gamma = u_j^H z
//                                                      (one pass over Q_j)
//   nu = sqrt(alpha - |s|^2),  q_j = (u_j - Q_j s)/nu   (second pass of u_j)
//   H[0:j, j-1] += s,  H[j, j-1] = nu                  (column j-1 complete)
//   p = H[0:j+1, 0:j] s                                 (A Q_j s = Q_{j+1} p)
//   h[0:j] = (t - p[0:j])/nu,  h[j] = ((gamma - s^H t)/nu - p[j])/nu
//                                                      (first pass of A q_j)
//   u_{j+1} = z/nu - (d/nu) u_j - Q_j (c - s d/nu),
//        c = p[0:j]/nu + h[0:j],  d = p[j]/nu + h[j]   (one pass over Q_j)
//
// Each step is three kernels over the N x M scratch T of the matvec:
//   A(j): pass 3 of z = A u_j (inverse row FFTs) + all inner products; the
//         last block completes column j-1, applies the FP64 Givens rotations
//         and prepares the coefficients of B(j).
//   B(j): q_j and u_{j+1} in one pass over Q_j + pass 1 of A u_{j+1}
//         (row FFTs of chi .* u_{j+1}).
//   pass 2 (column FFTs, spectrum multiply), shared with the plain matvec.
// Flexible steps (FGMRES with a right preconditioner P, possibly nonlinear):
// the network is applied to the lagged vector, z_j = P(u_j / ||u_j||), and
// w = A z_j takes the place of z = A u_j.  Any z_j keeps the flexible Arnoldi
// relation A Z_m = Q_{m+1} H, and the same single reduction gives the first
// pass of w against Q_{j+1} = [Q_j, q_j] directly:
//   h[0:j] = t,  h[j] = q_j^H w = (gamma - s^H t)/nu,
//   u_{j+1} = w - (h[j]/nu) u_j - Q_j (t - s h[j]/nu)
// (the plain form with z/nu -> w, c -> t, d -> h[j]).  B(j) then also reduces
// ||u_{j+1}||, the input scale of the next network application, and does no
// row FFTs: the next matvec input is z_{j+1}, known only after the network.
// The cycle's last step needs only s and alpha of u_m (column m-1 of H), so
// B(m-1) stops after u_m, pass 2 is not launched and A(m) skips pass 3: a
// cycle of m steps costs m matvecs.
// A block owns R whole rows (R = floor(1024*EPD / N); R*N = 1024*EPD for
// powers of two) so the row FFTs and the element-wise Krylov work run in the
// same kernel.
#include <math.h>

#include "cgm_fft.h"
#include "cgm_reduce.h"

namespace cgm {

constexpr int kFBS = 1024;
constexpr int kCH = 8;  // basis vectors per accumulation chunk

struct FusedArgs {
  int N, R, m, ld, j, mode;  // m: steps of the cycle; ld: leading dimension of H, Rm
  int tail;              // last launch of its kind in the cycle: no matvec work (A(m), B(m-1))
  int flex;              // flexible steps (see below)
  int G;                 // blocks
  long n;
  float inv_mm;          // 1 / M^2
  const float* chi;
  float2* V;             // q_0 .. q_m
  const float2* Uj;      // u_j
  const float2* Xin;     // matvec input: u_j, or z_j = P(u_j / ||u_j||) with flex
  double* partn;         // ||u_{j+1}||^2 partials (flex)
  float2* Unext;         // u_{j+1}
  float2* Z;             // z = A u_j
  float2* T;             // N x M FFT scratch
  const float2* tw;      // twiddles
  double2* part;         // partials: s [0, m+1), t [m+1, 2m+2), (alpha, gamma) [2m+2, 2m+4)
  double2* H;            // (m+1) x m Hessenberg, row-major, ld m
  double2* Rm;           // m x m rotated factor, row-major, ld m
  double* cs;
  double2* sn;
  double2* g;
  double2* h1p;          // first-pass coefficients of the previous step (m+1)
  double2* a1;           // coefficients of q_j:      Q_j s
  double2* a2;           // coefficients of u_{j+1}:  Q_j (c - s d/nu)
  double2* scal;         // [0] = (1/nu, 0), [1] = d/nu
  DevState* st;
  double* hist;          // estimates, hist[j] after j steps of this cycle (may be null)
  // per-system strides of a batched launch (system = blockIdx.y); elements,
  // DevState in bytes
  long sChi, sBasis, sU, sVec, sT, sPart, sSmall, sSt, sHist, sPartn, sXin;
};

__device__ inline void batch_offset(FusedArgs& a) {
  const long k = blockIdx.y;
  if (k == 0) return;
  a.chi += k * a.sChi;
  a.V += k * a.sBasis;
  a.Uj += k * a.sU;
  if (a.Xin) a.Xin += k * a.sXin;
  a.Unext += k * a.sU;
  if (a.partn) a.partn += k * a.sPartn;
  a.Z += k * a.sVec;
  a.T += k * a.sT;
  a.part += k * a.sPart;
  a.H += k * a.sSmall;
  a.Rm += k * a.sSmall;
  a.cs += k * a.sSmall;
  a.sn += k * a.sSmall;
  a.g += k * a.sSmall;
  a.h1p += k * a.sSmall;
  a.a1 += k * a.sSmall;
  a.a2 += k * a.sSmall;
  a.scal += k * a.sSmall;
  a.st = boff(a.st, k * a.sSt);
  if (a.hist) a.hist += k * a.sHist;
}

__device__ inline float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ inline float2 cjmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ inline double2 dmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 dcjmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ inline double2 dsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 dadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 dscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

template <int M, int EPD, int BS>
__global__ void __launch_bounds__(BS) k_arn_a(FusedArgs a) {
  batch_offset(a);
  const int dn = a.st->done;  // tested once this block's loads are in flight
  constexpr int NT = FftShape<M>::NT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  const int N = a.N, R = a.R, j = a.j;
  const long n = a.n;
  const int r0 = blockIdx.x * R;
  const int tid = threadIdx.x;
  // ---- pass 3: inverse row FFTs --------------------------------------------
  const bool fa = tid < R * NT;
  const int t = tid % NT, rl = fa ? tid / NT : 0;
  FftTwiddles<float, M> tws;
  if (fa) tws.load(a.tw, t);
  constexpr int ML = LdsLen<float, M>::value;
  float2* buf = smem + rl * ML;
  const bool tail = a.tail != 0;
  // the row values of pass 3 and this thread's u values, loaded before the
  // done test so that their latency overlaps it
  constexpr int EPT = FftShape<M>::EPT;
  float2 trow[EPT];
  if (!tail && fa) {
    const int row = r0 + rl;
#pragma unroll
    for (int k = 0; k < EPT; ++k)
      trow[k] = row < N ? a.T[tpos(row, t + k * NT, N)] : make_float2(0.f, 0.f);
  }
  // u_j and the matvec input (z_j for flexible steps, else u_j again: a
  // second independent load rather than a select that waits for the first)
  const float2* xsrc = a.flex && !tail ? a.Xin : a.Uj;
  float2 upre[EPD], xpre[EPD];
#pragma unroll
  for (int k = 0; k < EPD; ++k) {
    const int e = tid + k * BS;
    const int r = e / N, c = e - r * N;
    const bool in = r < R && r0 + r < N;
    const long id = in ? (long)(r0 + r) * N + c : 0;
    upre[k] = in ? a.Uj[id] : make_float2(0.f, 0.f);
    xpre[k] = in ? xsrc[id] : make_float2(0.f, 0.f);
  }
  if (dn) return;
  if (!tail) {
    if (fa) {
#pragma unroll
      for (int k = 0; k < EPT; ++k) buf[lpad<float>(t + k * NT)] = trow[k];
    }
    __syncthreads();
    lds_fft<float, M, +1, false, true>(buf, tws, t, fa);  // columns < N are used
  }
  // ---- z = u - G(.), u and z of the block's R*N elements staged in LDS -------
  // (the FFT rows are no longer needed once every thread has read its values)
  float2 uu[EPD], zz[EPD];
#pragma unroll
  for (int k = 0; k < EPD; ++k) {
    const int e = tid + k * BS;
    const int r = e / N, c = e - r * N;
    const int row = r0 + r;
    uu[k] = make_float2(0.f, 0.f);
    zz[k] = make_float2(0.f, 0.f);
    if (r < R && row < N) {
      const long id = (long)row * N + c;
      const float2 u = upre[k];
      uu[k] = u;
      if (!tail) {
        const float2 xin = xpre[k];
        float2 tv = smem[r * ML + lpad<float>(c)];
        tv.x *= a.inv_mm;
        tv.y *= a.inv_mm;
        if (a.mode == kModeAdjoint) {
          const float ch = a.chi[id];
          tv.x *= ch;
          tv.y *= ch;
        }
        const float2 z = make_float2(xin.x - tv.x, xin.y - tv.y);
        a.Z[id] = z;
        zz[k] = z;
      }
    }
  }
  __syncthreads();
  float2* su = smem;                 // R*N values of u
  float2* sz = smem + BS * EPD;    // R*N values of z
#pragma unroll
  for (int k = 0; k < EPD; ++k) {
    su[tid + k * BS] = uu[k];
    sz[tid + k * BS] = zz[k];
  }
  __syncthreads();
  // ---- inner products ----------------------------------------------------------
  // Work units (q, slice): basis vector q < j (slot q == j carries
  // alpha = u^H u and gamma = u^H z) over one slice of the block's elements
  // (a multiple of 256: batches of 8 loads per lane).  A vector is cut into
  // S = 32 / (j+1) slices (at least 1, at most one per 256 elements), so the
  // 32 lane groups all stream basis data at small j and each group reduces
  // once per vector at large j; each unit's 32-lane sums go to LDS and the S
  // slices of a vector are added in a fixed order (FP64) into the block
  // partial.
  // Partial layout (per launch): s_q / alpha at q, t_q / gamma at j+1+q.
  const int G = a.G;
  const long base = (long)r0 * N;  // the block's rows are contiguous
  const int nrows_live = (N - r0) < R ? (N - r0) : R;
  const int RNl = nrows_live * N;   // live elements of this block
  const int grp = tid >> 5, g = tid & 31;
  constexpr int kSlice = 256;
  __shared__ float4 lunit[65 * (BS * EPD / kSlice)];
  const int nb = (RNl + kSlice - 1) / kSlice;  // 256-element batches of the block
  int S = (BS / 32) / (j + 1);
  S = S < 1 ? 1 : (S > nb ? nb : S);
  const int slice = ((nb + S - 1) / S) * kSlice;
  const int units = (j + 1) * S;
  for (int u0 = 0; u0 < units; u0 += BS / 32) {
    const int u = u0 + grp;
    float2 as = make_float2(0.f, 0.f), at = make_float2(0.f, 0.f);
    if (u < units) {
      const int q = u / S;
      const int e_lo = (u - q * S) * slice;
      const int e_hi = (RNl < e_lo + slice) ? RNl : e_lo + slice;
      if (q < j) {
        const float2* Vq = a.V + (long)q * n + base;
        for (int e0 = e_lo + g; e0 < e_hi; e0 += 2 * kSlice) {
          float2 v[16];
#pragma unroll
          for (int k = 0; k < 16; ++k) {
            const int e = e0 + 32 * k;
            v[k] = e < e_hi ? Vq[e] : make_float2(0.f, 0.f);
          }
#pragma unroll
          for (int k = 0; k < 16; ++k) {
            const int e = e0 + 32 * k;
            if (e < e_hi) {
              const float2 ps = cjmulf(v[k], su[e]), pt = cjmulf(v[k], sz[e]);
              as.x += ps.x; as.y += ps.y;
              at.x += pt.x; at.y += pt.y;
            }
          }
        }
      } else {
        for (int e = e_lo + g; e < e_hi; e += 32) {
          const float2 uv = su[e], zv = sz[e];
          as.x += uv.x * uv.x + uv.y * uv.y;
          const float2 pg = cjmulf(uv, zv);
          at.x += pg.x; at.y += pg.y;
        }
      }
    }
    // 32-lane sums (rows 0+1 and rows 2+3 of the wave)
    const float sx = half_wave_sum_f(as.x), sy = half_wave_sum_f(as.y);
    const float tx = half_wave_sum_f(at.x), ty = half_wave_sum_f(at.y);
    if (g == 0 && u < units) lunit[u] = make_float4(sx, sy, tx, ty);
  }
  __syncthreads();
  for (int q = tid; q <= j; q += BS) {
    double sx = 0.0, sy = 0.0, tx = 0.0, ty = 0.0;
    for (int sl = 0; sl < S; ++sl) {
      const float4 pu = lunit[q * S + sl];
      sx += pu.x;
      sy += pu.y;
      tx += pu.z;
      ty += pu.w;
    }
    put2(&a.part[(long)q * G + blockIdx.x], sx, sy);
    put2(&a.part[(long)(j + 1 + q) * G + blockIdx.x], tx, ty);
  }
  if (!arrive_last(&a.st->tickets[kTicketDots1])) return;

  // ---- last block: FP64 small math ------------------------------------------
  constexpr int LM = 65;  // max restart + 1
  __shared__ double2 ls[LM], lt[LM], lp[LM + 1], lcol[LM + 1], lsn[LM];
  __shared__ double lcs[LM];
  __shared__ double2 lag[2];
  __shared__ double s_nu;
  __shared__ int s_stop;
  __shared__ double2 lall[2 * LM];
  __shared__ double2 lh1p[LM];
  __shared__ double2 s_g;
  DevState* st = a.st;
  // the small math's inputs, fetched together with the partial sums
  for (int i = tid; i < j; i += BS) {
    lh1p[i] = a.h1p[i];
    if (i < j - 1) {
      lcs[i] = a.cs[i];
      lsn[i] = a.sn[i];
    }
  }
  if (tid == 0 && j > 0) s_g = a.g[j - 1];
  finish_partials<BS>(a.part, G, 2 * (j + 1), lall);
  __syncthreads();
  for (int i = tid; i <= j; i += BS) {
    if (i < j) {
      ls[i] = lall[i];
      lt[i] = lall[j + 1 + i];
    } else {
      lag[0] = lall[j];
      lag[1] = lall[2 * j + 1];
    }
  }
  __syncthreads();
  if (tid == 0) {
    double ss = 0.0;
    for (int i = 0; i < j; ++i) ss += ls[i].x * ls[i].x + ls[i].y * ls[i].y;
    const double nu2 = lag[0].x - ss;
    const double nu = nu2 > 0.0 ? sqrt(nu2) : 0.0;
    s_nu = nu;
    int stop = 0;
    const bool bad = !isfinite(lag[0].x) || !isfinite(lag[1].x) || !isfinite(lag[1].y) || !isfinite(nu);
    if (bad) {
      st->status = CGM_ERROR_NONFINITE;
      stop = 1;
    }
    if (j == 0) {
      st->beta = nu;
      a.g[0] = make_double2(nu, 0.0);
      st->k = 0;
      if (!bad && !(nu > st->tol_abs)) stop = 1;
    }
    s_stop = stop;
  }
  __syncthreads();
  const double nu = s_nu;
  if (j > 0 && !s_stop) {
    // complete column j-1 of H and rotate it
    for (int i = tid; i <= j; i += BS) {
      double2 hv;
      if (i < j) hv = dadd(lh1p[i], ls[i]);
      else hv = make_double2(nu, 0.0);
      lcol[i] = hv;
      a.H[(long)i * a.ld + (j - 1)] = hv;
    }
    __syncthreads();
    if (tid == 0) {
      // rotated copy of the column in lp (lcol keeps the original for p below)
      for (int i = 0; i <= j; ++i) lp[i] = lcol[i];
      for (int i = 0; i < j - 1; ++i) {
        const double c = lcs[i];
        const double2 s = lsn[i];
        const double2 x0 = lp[i], x1 = lp[i + 1];
        lp[i] = make_double2(c * x0.x + (s.x * x1.x - s.y * x1.y), c * x0.y + (s.x * x1.y + s.y * x1.x));
        lp[i + 1] = make_double2(-(s.x * x0.x + s.y * x0.y) + c * x1.x, -(s.x * x0.y - s.y * x0.x) + c * x1.y);
      }
      const int i = j - 1;
      const double2 x0 = lp[i];
      const double aa = hypot(x0.x, x0.y);
      const double d = hypot(aa, nu);
      double c;
      double2 s, r;
      if (aa == 0.0) {
        c = 0.0;
        s = make_double2(1.0, 0.0);
        r = make_double2(nu, 0.0);
      } else {
        c = aa / d;
        s = make_double2(x0.x / aa * nu / d, x0.y / aa * nu / d);
        r = make_double2(x0.x / aa * d, x0.y / aa * d);
      }
      lp[i] = r;
      a.cs[i] = c;
      a.sn[i] = s;
      const double2 gi = s_g;  // g[j-1]
      const double2 gn = make_double2(-(s.x * gi.x + s.y * gi.y), -(s.x * gi.y - s.y * gi.x));
      a.g[i + 1] = gn;
      a.g[i] = make_double2(c * gi.x, c * gi.y);
      const double est = hypot(gn.x, gn.y);
      st->res_est = est;
      st->hn = nu;
      st->k = j;
      if (a.hist) a.hist[j] = est;
      if (!isfinite(est)) st->status = CGM_ERROR_NONFINITE;
      if (!isfinite(est) || est <= st->tol_abs || !(nu > 0.0) || j == a.m) s_stop = 1;
    }
    __syncthreads();
    for (int i = tid; i < j; i += BS) a.Rm[(long)i * a.ld + (j - 1)] = lp[i];
  }
  __syncthreads();
  if (s_stop || j == a.m) {
    if (tid == 0) {
      st->done = 1;
      reset_ticket(&st->tickets[kTicketDots1]);
    }
    return;
  }
  if (a.flex) {
    // first pass of w against Q_{j+1} and the update coefficients
    __shared__ double2 s_hj;
    if (tid == 0) {
      double2 st_ = make_double2(0.0, 0.0);
      for (int i = 0; i < j; ++i) st_ = dadd(st_, dcjmul(ls[i], lt[i]));
      const double2 hj = dscale(dsub(lag[1], st_), 1.0 / nu);
      a.h1p[j] = hj;
      s_hj = hj;
      a.scal[0] = make_double2(1.0 / nu, 0.0);
      a.scal[1] = dscale(hj, 1.0 / nu);
      a.scal[2] = make_double2(1.0, 0.0);
    }
    __syncthreads();
    const double2 dnu = dscale(s_hj, 1.0 / nu);
    for (int i = tid; i < j; i += BS) {
      a.h1p[i] = lt[i];
      a.a1[i] = ls[i];
      a.a2[i] = dsub(lt[i], dmul(ls[i], dnu));
    }
    if (tid == 0) reset_ticket(&st->tickets[kTicketDots1]);
    return;
  }
  // p = H[0:j+1, 0:j] s: one wave per row, lane c multiplies column c
  {
    const int wave = tid >> 6, lane = tid & 63;
    for (int r = wave; r <= j; r += BS / 64) {
      double2 v = make_double2(0.0, 0.0);
      if (lane < j && lane >= r - 1) {
        const double2 h = (lane == j - 1) ? lcol[r] : a.H[(long)r * a.ld + lane];
        v = dmul(h, ls[lane]);
      }
      v.x = wave_sum(v.x);
      v.y = wave_sum(v.y);
      if (lane == 0) lp[r] = v;
    }
  }
  __syncthreads();
  // first-pass coefficients of A q_j and the update coefficients
  __shared__ double2 s_dd;
  if (tid == 0) {
    double2 st_ = make_double2(0.0, 0.0);
    for (int i = 0; i < j; ++i) st_ = dadd(st_, dcjmul(ls[i], lt[i]));
    const double2 qz = dscale(dsub(lag[1], st_), 1.0 / nu);
    const double2 hj = dscale(dsub(qz, lp[j]), 1.0 / nu);
    a.h1p[j] = hj;
    const double2 dd = dadd(dscale(lp[j], 1.0 / nu), hj);
    s_dd = dd;
    a.scal[0] = make_double2(1.0 / nu, 0.0);
    a.scal[1] = dscale(dd, 1.0 / nu);
    a.scal[2] = make_double2(1.0 / nu, 0.0);
  }
  __syncthreads();
  const double2 ddnu = dscale(s_dd, 1.0 / nu);
  for (int i = tid; i < j; i += BS) {
    const double2 hi = dscale(dsub(lt[i], lp[i]), 1.0 / nu);
    a.h1p[i] = hi;
    const double2 ci = dadd(dscale(lp[i], 1.0 / nu), hi);
    a.a1[i] = ls[i];
    a.a2[i] = dsub(ci, dmul(ls[i], ddnu));
  }
  if (tid == 0) reset_ticket(&st->tickets[kTicketDots1]);
}

template <int M, int EPD, int BS>
__global__ void __launch_bounds__(BS) k_arn_b(FusedArgs a) {
  batch_offset(a);
  const int dn = a.st->done;  // tested once this block's u, z and basis loads are issued
  constexpr int NT = FftShape<M>::NT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float2* smem = reinterpret_cast<float2*>(smem_raw);
  const int N = a.N, R = a.R, j = a.j;
  const long n = a.n;
  const int r0 = blockIdx.x * R;
  const int tid = threadIdx.x;
  const bool fa = tid < R * NT;
  const int t = tid % NT, rl = fa ? tid / NT : 0;
  FftTwiddles<float, M> tws;
  if (fa) tws.load(a.tw, t);
  const double2 sc0 = a.scal[0], sc1 = a.scal[1];
  const float invnu = (float)sc0.x;
  const float wsc = (float)a.scal[2].x;  // coefficient of z (plain: 1/nu) or w (flex: 1)
  const float2 ddnu = make_float2((float)sc1.x, (float)sc1.y);
  float uu2 = 0.f;
  float2 up[EPD], zp[EPD];
#pragma unroll
  for (int k = 0; k < EPD; ++k) {
    const int e = tid + k * BS;
    const int r = e / N, c = e - r * N;
    const bool live = r < R && r0 + r < N;
    const long id = (long)(r0 + r) * N + c;
    up[k] = live ? a.Uj[id] : make_float2(0.f, 0.f);
    zp[k] = live ? a.Z[id] : make_float2(0.f, 0.f);
  }
  // Q_j a1 and Q_j a2 for the thread's EPD elements, QC basis vectors at a
  // time for all elements (QC * EPD loads in flight)
  constexpr int QC = EPD == 1 ? 4 : 2;
  float2 qa[EPD], qb[EPD];
  long ids[EPD];
  bool lv[EPD];
#pragma unroll
  for (int k = 0; k < EPD; ++k) {
    const int e = tid + k * BS;
    const int r = e / N, c = e - r * N;
    lv[k] = r < R && r0 + r < N;
    ids[k] = lv[k] ? (long)(r0 + r) * N + c : 0;
    qa[k] = make_float2(0.f, 0.f);
    qb[k] = make_float2(0.f, 0.f);
  }
  {
    int i = 0;
    for (; i + QC <= j; i += QC) {
      float2 v[QC][EPD];
#pragma unroll
      for (int q = 0; q < QC; ++q)
#pragma unroll
        for (int k = 0; k < EPD; ++k) v[q][k] = lv[k] ? a.V[(long)(i + q) * n + ids[k]] : make_float2(0.f, 0.f);
#pragma unroll
      for (int q = 0; q < QC; ++q) {
        const double2 c1 = a.a1[i + q], c2 = a.a2[i + q];
        const float2 f1 = make_float2((float)c1.x, (float)c1.y), f2 = make_float2((float)c2.x, (float)c2.y);
#pragma unroll
        for (int k = 0; k < EPD; ++k) {
          const float2 p1 = cmulf(f1, v[q][k]), p2 = cmulf(f2, v[q][k]);
          qa[k].x += p1.x; qa[k].y += p1.y;
          qb[k].x += p2.x; qb[k].y += p2.y;
        }
      }
    }
    for (; i < j; ++i) {
      const double2 c1 = a.a1[i], c2 = a.a2[i];
      const float2 f1 = make_float2((float)c1.x, (float)c1.y), f2 = make_float2((float)c2.x, (float)c2.y);
#pragma unroll
      for (int k = 0; k < EPD; ++k) {
        const float2 v = lv[k] ? a.V[(long)i * n + ids[k]] : make_float2(0.f, 0.f);
        const float2 p1 = cmulf(f1, v), p2 = cmulf(f2, v);
        qa[k].x += p1.x; qa[k].y += p1.y;
        qb[k].x += p2.x; qb[k].y += p2.y;
      }
    }
  }
  // the done test comes after the basis loads (a finished system's block
  // reads in-bounds stale columns and writes nothing), so that waiting for
  // it does not also wait for u and z before the first basis load is issued
  if (dn) return;
#pragma unroll
  for (int k = 0; k < EPD; ++k) {
    const int e = tid + k * BS;
    const int r = e / N, c = e - r * N;
    float2 w = make_float2(0.f, 0.f);
    if (lv[k]) {
      const long id = ids[k];
      const float2 u = up[k], z = zp[k];
      const float2 q = make_float2((u.x - qa[k].x) * invnu, (u.y - qa[k].y) * invnu);
      a.V[(long)j * n + id] = q;
      const float2 du = cmulf(ddnu, u);
      const float2 un = make_float2(z.x * wsc - du.x - qb[k].x, z.y * wsc - du.y - qb[k].y);
      a.Unext[id] = un;
      uu2 += un.x * un.x + un.y * un.y;
      w = un;
      if (a.mode == kModeForward && !a.flex) {  // flexible steps transform z_j, not u_{j+1}
        const float ch = a.chi[id];
        w.x *= ch;
        w.y *= ch;
      }
    }
    if (r < R && !a.flex) smem[r * LdsLen<float, M>::value + lpad<float>(c)] = w;
  }
  if (a.flex) {
    // ||u_{j+1}||: block partial, the last block sums them in a fixed order
    if (a.tail) return;  // no network application follows u_m
    __shared__ double red[BS / 64];
    const double ws_ = wave_sum((double)uu2);
    if ((tid & 63) == 0) red[tid >> 6] = ws_;
    __syncthreads();
    if (tid == 0) {
      double sb = 0.0;
      for (int k = 0; k < BS / 64; ++k) sb += red[k];
      put(&a.partn[blockIdx.x], sb);
    }
    if (!arrive_last(&a.st->tickets[kTicketNorm])) return;
    const double sq = sum_partials(a.partn, a.G);
    if (tid == 0) {
      a.st->unorm = sqrt(sq);
      reset_ticket(&a.st->tickets[kTicketNorm]);
    }
    return;
  }
  if (a.tail) return;  // u_m needs no matvec (block-uniform)
  constexpr int ZE = fft_lo_read_end<M>();  // zeros the transform reads
  for (int e = tid; e < R * (ZE - N); e += BS) {
    const int r = e / (ZE - N), c = N + e % (ZE - N);
    smem[r * LdsLen<float, M>::value + lpad<float>(c)] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2* buf = smem + rl * LdsLen<float, M>::value;
  lds_fft<float, M, -1, true>(buf, tws, t, fa);  // the rows are zero from N <= M/2 on
  if (fa) {
    const int row = r0 + rl;
    if (row < N) {
      for (int e = t; e < M; e += NT) a.T[tpos(row, e, N)] = buf[lpad<float>(e)];
    }
  }
}

// ---------------------------------------------------------------------------
bool fused_supported(int N, int restart) {
  const bool pow2 = N >= 64 && N <= 2048 && (N & (N - 1)) == 0;
  const bool three = N == 96 || N == 192 || N == 384 || N == 768;
  return (pow2 || three) && restart >= 1 && restart <= 64;
}

// Work-group shapes.  Large: 1,024 threads, one element per thread (two from
// N = 1024 and at N = 384), so a block holds at least two rows where
// registers allow.  Small (N in {64, 96, 128, 192, 256, 512}): 256 threads
// with four elements each; the same rows per block, so every wave of the
// block takes part in the row transforms and four blocks share a CU.
struct FusedShape {
  int epd, bs;
};
static bool fused_small_ok(int N) {
  return N == 64 || N == 96 || N == 128 || N == 192 || N == 256 || N == 512;
}
static FusedShape fused_shape(int N, bool small) {
  if (small && fused_small_ok(N)) return {4, 256};
  return {N >= 1024 || N == 384 ? 2 : 1, kFBS};
}

int fused_blocks(int N) {
  const FusedShape f = fused_shape(N, false);  // both shapes give the same rows per block
  const int R = f.bs * f.epd / N;
  return (N + R - 1) / R;
}

// One launch of A(j) (which = 0) or B(j) (which = 1).
int launch_fused(hipStream_t st, int which, const FusedLaunch& L) {
  FusedArgs a;
  a.N = L.N;
  const int M = 2 * L.N;
  const FusedShape fs = fused_shape(L.N, L.small != 0);
  const int epd = fs.epd;
  a.R = fs.bs * epd / L.N;
  a.m = L.m;
  a.ld = L.ld;
  a.j = L.j;
  a.mode = L.mode;
  a.flex = L.flex;
  a.Xin = L.Xin;
  a.partn = L.partn;
  a.sPartn = L.sPartn;
  a.sXin = L.sXin;
  a.tail = which == 0 ? (L.j == L.m) : (L.j == L.m - 1);
  a.G = fused_blocks(L.N);
  a.n = (long)L.N * L.N;
  a.inv_mm = (float)(1.0 / ((double)M * (double)M));
  a.chi = L.chi;
  a.V = L.V;
  a.Uj = L.Uj;
  a.Unext = L.Unext;
  a.Z = L.Z;
  a.T = L.T;
  a.tw = L.tw;
  a.part = L.part;
  a.H = L.H;
  a.Rm = L.Rm;
  a.cs = L.cs;
  a.sn = L.sn;
  a.g = L.g;
  a.h1p = L.h1p;
  a.a1 = L.a1;
  a.a2 = L.a2;
  a.scal = L.scal;
  a.st = L.st;
  a.hist = L.hist;
  a.sChi = L.sChi;
  a.sBasis = L.sBasis;
  a.sU = L.sU;
  a.sVec = L.sVec;
  a.sT = L.sT;
  a.sPart = L.sPart;
  a.sSmall = L.sSmall;
  a.sSt = L.sSt;
  a.sHist = L.sHist;
  // the row FFTs, then u and z of the block staged in the same space
  size_t lds = (size_t)a.R * lds_len(M, false) * sizeof(float2);
  const size_t stage = (size_t)2 * fs.bs * epd * sizeof(float2);
  if (lds < stage) lds = stage;
#define CGM_FUSED_LAUNCH(NN, EPD, BSZ)                                                          \
  {                                                                                           \
    if (which == 0)                                                                           \
      hipLaunchKernelGGL((k_arn_a<2 * NN, EPD, BSZ>), dim3(a.G, L.B), dim3(BSZ), lds, st, a); \
    else                                                                                      \
      hipLaunchKernelGGL((k_arn_b<2 * NN, EPD, BSZ>), dim3(a.G, L.B), dim3(BSZ), lds, st, a); \
  }
#define CGM_FUSED_CASE(NN, EPD)                                                                 \
  case NN:                                                                                    \
    CGM_FUSED_LAUNCH(NN, EPD, kFBS)                                                           \
    break;
#define CGM_FUSED_CASE2(NN, EPD)                                                                \
  case NN:                                                                                    \
    if (fs.bs == 256) CGM_FUSED_LAUNCH(NN, 4, 256)                                            \
    else CGM_FUSED_LAUNCH(NN, EPD, kFBS)                                                      \
    break;
  switch (L.N) {
    CGM_FUSED_CASE2(64, 1)
    CGM_FUSED_CASE2(128, 1)
    CGM_FUSED_CASE2(256, 1)
    CGM_FUSED_CASE2(512, 1)
    CGM_FUSED_CASE(1024, 2)
    CGM_FUSED_CASE(2048, 2)
    CGM_FUSED_CASE2(96, 1)
    CGM_FUSED_CASE2(192, 1)
    CGM_FUSED_CASE(384, 2)
    CGM_FUSED_CASE(768, 1)
#undef CGM_FUSED_CASE
#undef CGM_FUSED_CASE2
#undef CGM_FUSED_LAUNCH
    default:
      return CGM_ERROR_UNSUPPORTED;
  }
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

}  // namespace cgm
