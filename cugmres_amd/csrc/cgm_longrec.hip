// Long-recurrence flexible GMRES: restart lengths up to kMaxRestartLR with
// the whole Krylov basis resident in HBM.
//
// On the high-contrast benchmark maps a restart of 30-256 steps discards the
// directions that took hundreds of steps to find, and restarted GMRES stalls
// (FGMRES(256) with the tile preconditioner ends at 1e-2 at N = 256); a
// 288 GB GPU holds thousands of basis vectors per system (2 MiB each at
// N = 512), so the cycle can be as long as the solve.  At that length the
// step cost is the basis traffic, and this path is built for it:
//   - lagged CGS2 (the DCGS2 scheme of cgm_arnoldi.hip, flexible form): one
//     projection pass and one update pass over the basis per step, so the
//     basis is read twice per step (the two-pass CGS2 kernels read it four
//     times once it has more than 32 vectors);
//   - the projection kernel tiles the basis 16 vectors x 1024 elements per
//     work-group (u_j and w in registers, re-read once per 16 vectors), the
//     update kernel streams one element per thread with 16 basis loads in
//     flight, so both have tens of thousands of loads outstanding at large j;
//   - partial sums are reduced in a fixed order by a separate kernel (no
//     tickets), the FP64 Givens update of the completed column runs in one
//     work-group per system with the rotations staged through LDS.
// Step j (Q_j = [q_0..q_{j-1}] orthonormal, u_j the once-orthogonalised
// vector, ||u_j|| known):
//   z_j = P(u_j / ||u_j||),  w = A z_j                  (preconditioner, matvec)
//   s = Q_j^H u_j, t = Q_j^H w, alpha = u_j^H u_j, gamma = u_j^H w   (PROJ, RED)
//   nu = sqrt(alpha - |s|^2), column j-1 of H = [h1p + s; nu], Givens (FIN)
//   h1p = [t; (gamma - s^H t)/nu], c = h1p[j]/nu
//   q_j = (u_j - Q_j s)/nu,  u_{j+1} = w - c u_j - Q_j (t - s c)      (UPD)
// (the flexible form of the lagged scheme: A Z_m = Q_{m+1} H holds for any
// z_j).  Step m runs PROJ/RED/FIN on u_m only, completing column m-1.
// Systems of a batch (grid y) are independent; each has its own basis and
// state and stops at its own tolerance.
#include <math.h>

#include <vector>

#include "cgm_reduce.h"

namespace cgm {

namespace {

constexpr int kBS = 256;     // threads per streaming work-group
constexpr int kPE = 16;      // elements per thread in the projection (chunk = 4096)
constexpr int kPP = kPE / 2; // element pairs per thread (one 16-byte load each)
constexpr int kJT = 16;      // basis vectors per projection work-group
constexpr int kNV = 2 * kJT + 2;  // values per projection work-group: s, t, (alpha, gamma)
constexpr int kUU = 16;      // basis loads in flight per thread in the update
constexpr int kStage = 1024; // rotations / coefficients staged per LDS round

// Per-system state of the long-recurrence path (device), at a stride of
// `small` elements per system for the FP64 vectors and `ld*ld` for R.
struct LrArgs {
  int j, m, ld, nchunks, ngroups, nuchunks;
  long n;
  float2* Q;         // [B][ld+1][n] basis
  float2* Z;         // [B][ld][n] preconditioned vectors
  const float2* U;   // u_j (stride sU)
  float2* Unext;     // u_{j+1}
  long sU;
  const float2* W;   // w = A z_j (stride sW)
  long sW;
  const double2* W64;  // the same in FP64 (stride n) when the products are FP64, else null
  double2* part;     // [B][ngroups_max][kNV][nchunks]
  long sPart;
  double2* s;        // [B][ld+1]
  double2* t;
  double2* h1p;
  double2* g;
  double2* y;
  double* cs;
  double2* sn;
  double2* a1;       // [B][ld+1] update coefficients
  double2* a2;
  double* scal;      // [B][8]: alpha, gamma, 1/nu, c.re, c.im
  long sSmall;
  double2* R;        // [B][ld][ld] column-major
  double* partn;     // [B][nuchunks] norm partials of u_{j+1}
  long sPartn;
  DevState* st;
  long sst;
  double* hist;      // [B][hstride] (may be null): entry j = estimate after j steps
  long shist;
};

__device__ inline bool sys_done(const LrArgs& a, const DevState*& st) {
  st = boff(a.st, (long)blockIdx.y * a.sst);
  return st->done != 0;
}

__device__ inline float2 cj_mul(float2 a, float2 b) {  // conj(a) b
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Streaming loads of the Krylov basis: each basis value is read once per
// pass and not again before the next step, so the loads carry the
// non-temporal hint.
typedef float f2v_t __attribute__((ext_vector_type(2)));
typedef float f4v_t __attribute__((ext_vector_type(4)));
__device__ inline float2 nt_load(const float2* p) {
  const f2v_t v = __builtin_nontemporal_load(reinterpret_cast<const f2v_t*>(p));
  return make_float2(v.x, v.y);
}
__device__ inline float4 nt_load(const float4* p) {
  const f4v_t v = __builtin_nontemporal_load(reinterpret_cast<const f4v_t*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

// Elements 2p and 2p+1 of a complex vector of n elements: one 16-byte load
// when n is even (every row of the basis then starts 16-byte aligned), two
// 8-byte loads otherwise; zero beyond n.
template <bool kVec, bool kNT = false>
__device__ inline void load_pair(const float2* __restrict__ v, long p, long n, float2& lo, float2& hi) {
  const long e = 2 * p;
  if (kVec) {
    if (e < n) {
      const float4 q = kNT ? nt_load(reinterpret_cast<const float4*>(v + e)) : *reinterpret_cast<const float4*>(v + e);
      lo = make_float2(q.x, q.y);
      hi = make_float2(q.z, q.w);
    } else {
      lo = hi = make_float2(0.f, 0.f);
    }
  } else {
    lo = e < n ? (kNT ? nt_load(v + e) : v[e]) : make_float2(0.f, 0.f);
    hi = e + 1 < n ? (kNT ? nt_load(v + e + 1) : v[e + 1]) : make_float2(0.f, 0.f);
  }
}

// PROJ: partial sums of s_v = q_v^H u, t_v = q_v^H w for the kJT vectors of
// group g over chunk c (4096 elements, 16 per thread), and (group 0)
// alpha = u^H u, gamma = u^H w.  Stored value-major,
// part[sys][g][value][chunk], so the reduction reads contiguous runs.
// with_w = 0 at the cycle's last step (u_m only).
//
// Precision: each thread forms FP32 sums of four terms and accumulates them
// in FP64; the work-group sums are FP64 (the long recurrences need inner
// products accurate to the FP32 rounding of each term).  The 32 FP64 sums
// of eight vectors (s, t; re, im) are reduced across a wave by a transposing
// butterfly: at each of five exchange steps a lane keeps half of its values
// and adds the partner lane's copy of that half, so lane L ends with value
// L / 2 (one exchange per value instead of six).
template <bool kVec>
__global__ void __launch_bounds__(kBS) k_lr_proj(LrArgs a, int with_w) {
  const DevState* st;
  if (sys_done(a, st)) return;
  const int y = blockIdx.y;
  const int g = blockIdx.x / a.nchunks, c = blockIdx.x % a.nchunks;
  const int v0 = g * kJT;
  const int nv = min(kJT, a.j - v0);
  const long pbase = (long)c * (kBS * kPP) + threadIdx.x;
  const float2* u = a.U + (long)y * a.sU;
  const float2* w = a.W + (long)y * a.sW;
  const float2* Q = a.Q + (long)y * (a.ld + 1) * a.n;
  // w in FP64: the FP64 product when there is one (rounding A z_j to
  // complex64 would put an error of the size of ||A z_j|| times the
  // complex64 unit roundoff into every column of the Arnoldi relation), else
  // the complex64 w widened
  float2 ue[kPE];
  double2 we[kPE];
  const double2* w64 = a.W64 ? a.W64 + (long)y * a.n : nullptr;
#pragma unroll
  for (int k = 0; k < kPP; ++k) {
    load_pair<kVec>(u, pbase + (long)k * kBS, a.n, ue[2 * k], ue[2 * k + 1]);
    const long e = 2 * (pbase + (long)k * kBS);
    if (with_w && w64) {
      we[2 * k] = e < a.n ? w64[e] : make_double2(0.0, 0.0);
      we[2 * k + 1] = e + 1 < a.n ? w64[e + 1] : make_double2(0.0, 0.0);
    } else if (with_w) {
      float2 lo, hi;
      load_pair<kVec>(w, pbase + (long)k * kBS, a.n, lo, hi);
      we[2 * k] = make_double2(lo.x, lo.y);
      we[2 * k + 1] = make_double2(hi.x, hi.y);
    } else {
      we[2 * k] = we[2 * k + 1] = make_double2(0.0, 0.0);
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __shared__ double red[kBS / 64][32];
  double* P = reinterpret_cast<double*>(a.part + (long)y * a.sPart + (long)g * kNV * a.nchunks + c);
#pragma unroll
  for (int h = 0; h < kJT / 8; ++h) {
    double acc[32];  // [vector 0..7][s.re, s.im, t.re, t.im]
#pragma unroll
    for (int i = 0; i < 32; ++i) acc[i] = 0.0;
#pragma unroll
    for (int vv = 0; vv < 8; ++vv) {
      const int v = h * 8 + vv;
      if (v < nv) {
        const float2* q = Q + (long)(v0 + v) * a.n;
        float2 qe[kPE];
#pragma unroll
        for (int k = 0; k < kPP; ++k) load_pair<kVec, true>(q, pbase + (long)k * kBS, a.n, qe[2 * k], qe[2 * k + 1]);
#pragma unroll
        for (int b4 = 0; b4 < kPE; b4 += 4) {
          float sr = 0.f, si = 0.f;
          double tr = 0.0, ti = 0.0;
#pragma unroll
          for (int k = b4; k < b4 + 4; ++k) {
            const float2 p = cj_mul(qe[k], ue[k]);
            sr += p.x;
            si += p.y;
            const double qx = qe[k].x, qy = qe[k].y;
            tr += qx * we[k].x + qy * we[k].y;
            ti += qx * we[k].y - qy * we[k].x;
          }
          acc[vv * 4 + 0] += sr;
          acc[vv * 4 + 1] += si;
          acc[vv * 4 + 2] += tr;
          acc[vv * 4 + 3] += ti;
        }
      }
    }
#pragma unroll
    for (int stp = 0; stp < 5; ++stp) {
      const int off = 32 >> stp, half = 16 >> stp;
      const bool upper = (lane & off) != 0;
#pragma unroll
      for (int i = 0; i < half; ++i) {
        const double keep = upper ? acc[i + half] : acc[i];
        const double send = upper ? acc[i] : acc[i + half];
        acc[i] = keep + __shfl_xor(send, off, 64);
      }
    }
    const double tot = acc[0] + __shfl_xor(acc[0], 1, 64);
    if ((lane & 1) == 0) red[wave][lane >> 1] = tot;
    __syncthreads();
    if (threadIdx.x < 32) {
      const int vv = threadIdx.x >> 2, comp = threadIdx.x & 3, v = h * 8 + vv;
      if (v < nv) {
        double sum = 0.0;
#pragma unroll
        for (int wv = 0; wv < kBS / 64; ++wv) sum += red[wv][threadIdx.x];
        const int val = (comp < 2 ? v : kJT + v);
        P[2 * (long)val * a.nchunks + (comp & 1)] = sum;
      }
    }
    __syncthreads();
  }
  if (g == 0) {
    double al = 0.0, gr = 0.0, gi = 0.0;
#pragma unroll
    for (int b4 = 0; b4 < kPE; b4 += 4) {
      float a4 = 0.f;
#pragma unroll
      for (int k = b4; k < b4 + 4; ++k) {
        a4 += ue[k].x * ue[k].x + ue[k].y * ue[k].y;
        const double ux = ue[k].x, uy = ue[k].y;
        gr += ux * we[k].x + uy * we[k].y;
        gi += ux * we[k].y - uy * we[k].x;
      }
      al += a4;
    }
    al = wave_sum(al);
    gr = wave_sum(gr);
    gi = wave_sum(gi);
    if (lane == 0) {
      red[wave][0] = al;
      red[wave][1] = gr;
      red[wave][2] = gi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
      for (int wv = 0; wv < kBS / 64; ++wv) {
        s0 += red[wv][0];
        s1 += red[wv][1];
        s2 += red[wv][2];
      }
      double2* Pd = reinterpret_cast<double2*>(P);
      Pd[(long)(2 * kJT) * a.nchunks] = make_double2(s0, 0.0);
      Pd[(long)(2 * kJT + 1) * a.nchunks] = make_double2(s1, s2);
    }
  }
}

// RED: sums of the chunk partials in a fixed order; group g writes s, t of
// its vectors and (g = 0) alpha, gamma.
__global__ void __launch_bounds__(kBS) k_lr_red(LrArgs a) {
  const DevState* st;
  if (sys_done(a, st)) return;
  const int y = blockIdx.y;
  const int g = blockIdx.x;
  const int v0 = g * kJT;
  const int nv = min(kJT, a.j - v0);
  __shared__ double2 out[kNV];
  // values of this group: s (nv), t (nv), alpha/gamma (group 0)
  const double2* P = a.part + (long)y * a.sPart + (long)g * kNV * a.nchunks;
  for (int idx = threadIdx.x / 4; idx < kNV; idx += kBS / 4) {
    const bool valid = (idx < kJT && idx < nv) || (idx >= kJT && idx < 2 * kJT && idx - kJT < nv) ||
                       (idx >= 2 * kJT && g == 0);
    double re = 0.0, im = 0.0;
    if (valid) {
      for (int c = threadIdx.x % 4; c < a.nchunks; c += 4) {
        const double2 p = P[(long)idx * a.nchunks + c];
        re += p.x;
        im += p.y;
      }
    }
    re += __shfl_xor(re, 1, 64);
    im += __shfl_xor(im, 1, 64);
    re += __shfl_xor(re, 2, 64);
    im += __shfl_xor(im, 2, 64);
    if (threadIdx.x % 4 == 0) out[idx] = make_double2(re, im);
  }
  __syncthreads();
  double2* s = a.s + (long)y * a.sSmall;
  double2* t = a.t + (long)y * a.sSmall;
  if ((int)threadIdx.x < nv) {
    s[v0 + threadIdx.x] = out[threadIdx.x];
    t[v0 + threadIdx.x] = out[kJT + threadIdx.x];
  }
  if (g == 0 && threadIdx.x == 0) {
    double* sc = a.scal + (long)y * 8;
    sc[0] = out[2 * kJT].x;
    sc[1] = out[2 * kJT + 1].x;
    sc[2] = out[2 * kJT + 1].y;
  }
}

__device__ inline double2 dmul2(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 dcj2(double2 a, double2 b) { return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }

// FIN: nu, column j-1 of H with its Givens rotations, the estimate, and the
// coefficients of UPD.  One work-group per system.
__global__ void __launch_bounds__(kBS) k_lr_fin(LrArgs a) {
  DevState* st = boff(a.st, (long)blockIdx.y * a.sst);
  if (st->done) return;
  const int y = blockIdx.y;
  const int j = a.j;
  const double2* s = a.s + (long)y * a.sSmall;
  const double2* t = a.t + (long)y * a.sSmall;
  double2* h1p = a.h1p + (long)y * a.sSmall;
  double2* gv = a.g + (long)y * a.sSmall;
  double* cs = a.cs + (long)y * a.sSmall;
  double2* sn = a.sn + (long)y * a.sSmall;
  double2* a1 = a.a1 + (long)y * a.sSmall;
  double2* a2 = a.a2 + (long)y * a.sSmall;
  double* sc = a.scal + (long)y * 8;
  double2* Rc = a.R + (long)y * a.ld * a.ld + (long)(j > 0 ? j - 1 : 0) * a.ld;  // column j-1
  // |s|^2 and s^H t
  double ss = 0.0, str = 0.0, sti = 0.0;
  for (int i = threadIdx.x; i < j; i += kBS) {
    const double2 si = s[i], ti = t[i];
    ss += si.x * si.x + si.y * si.y;
    const double2 p = dcj2(si, ti);
    str += p.x;
    sti += p.y;
  }
  __shared__ double red[3][kBS / 64];
  ss = wave_sum(ss);
  str = wave_sum(str);
  sti = wave_sum(sti);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = ss;
    red[1][threadIdx.x >> 6] = str;
    red[2][threadIdx.x >> 6] = sti;
  }
  __syncthreads();
  ss = str = sti = 0.0;
  for (int w = 0; w < kBS / 64; ++w) {
    ss += red[0][w];
    str += red[1][w];
    sti += red[2][w];
  }
  const double alpha = sc[0];
  const double nu2 = alpha - ss;
  const double nu = nu2 > 0.0 ? sqrt(nu2) : 0.0;
  __shared__ int s_done;
  if (threadIdx.x == 0) s_done = 0;
  if (j > 0) {
    // column j-1: rows 0..j-1 = h1p + s, row j = nu, then rotations 0..j-2
    for (int i = threadIdx.x; i < j; i += kBS) Rc[i] = make_double2(h1p[i].x + s[i].x, h1p[i].y + s[i].y);
    __syncthreads();
    // The previous rotations 0..j-2 act on the pairs (i, i+1) in turn; the
    // entry carried down the column obeys carry_{i+1} = f_i(carry_i) with
    // the affine map f_i(a) = -conj(s_i) a + c_i col_{i+1}, and the rotated
    // row i is c_i carry_i + s_i col_{i+1}.  Affine maps compose
    // associatively, so the chain is a block-wide scan: each thread composes
    // the maps of its contiguous segment, the segments' maps are scanned in
    // LDS (Kogge-Stone), and each thread then replays its segment from its
    // incoming carry.  O(j / 256 + log 256) steps instead of j.
    __shared__ double2 scA[2][kBS], scB[2][kBS];
    __shared__ double2 carry;
    const int nrot = j - 1;
    const int per = (nrot + kBS - 1) / kBS;
    const int r0 = min(nrot, (int)threadIdx.x * per), r1 = min(nrot, r0 + per);
    const double2 carry0 = Rc[0];
    double2 A = make_double2(1.0, 0.0), Bm = make_double2(0.0, 0.0);
    for (int i = r0; i < r1; ++i) {
      const double c = cs[i];
      const double2 sv = sn[i], col = Rc[i + 1];
      const double2 al = make_double2(-sv.x, sv.y);   // -conj(s_i)
      A = dmul2(al, A);
      Bm = dmul2(al, Bm);
      Bm.x += c * col.x;
      Bm.y += c * col.y;
    }
    // the first entry of the next segment, which its thread overwrites below
    const double2 bnd = (r1 > r0) ? Rc[r1] : make_double2(0.0, 0.0);
    int buf = 0;
    scA[0][threadIdx.x] = A;
    scB[0][threadIdx.x] = Bm;
    __syncthreads();
    for (int d = 1; d < kBS; d <<= 1) {
      double2 a2 = scA[buf][threadIdx.x], b2 = scB[buf][threadIdx.x];
      if ((int)threadIdx.x >= d) {
        // (this segment's prefix) after (the prefix d segments earlier)
        const double2 ae = scA[buf][threadIdx.x - d], be = scB[buf][threadIdx.x - d];
        const double2 nb = dmul2(a2, be);
        b2 = make_double2(nb.x + b2.x, nb.y + b2.y);
        a2 = dmul2(a2, ae);
      }
      scA[buf ^ 1][threadIdx.x] = a2;
      scB[buf ^ 1][threadIdx.x] = b2;
      buf ^= 1;
      __syncthreads();
    }
    double2 av = carry0;
    if (threadIdx.x > 0) {
      const double2 ae = scA[buf][threadIdx.x - 1], be = scB[buf][threadIdx.x - 1];
      const double2 t = dmul2(ae, carry0);
      av = make_double2(t.x + be.x, t.y + be.y);
    }
    for (int i = r0; i < r1; ++i) {
      const double c = cs[i];
      const double2 sv = sn[i], col = (i + 1 < r1) ? Rc[i + 1] : bnd;
      const double2 sc2 = dmul2(sv, col);
      const double2 al = make_double2(-sv.x, sv.y);
      const double2 nb = dmul2(al, av);
      Rc[i] = make_double2(c * av.x + sc2.x, c * av.y + sc2.y);
      av = make_double2(nb.x + c * col.x, nb.y + c * col.y);
    }
    if (r1 == nrot && r1 > r0) carry = av;
    if (nrot == 0 && threadIdx.x == 0) carry = carry0;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double2 av = carry;
      const double aa = hypot(av.x, av.y);
      const double d = hypot(aa, nu);
      double c;
      double2 sv, r;
      if (aa == 0.0) {
        c = 0.0;
        sv = make_double2(1.0, 0.0);
        r = make_double2(nu, 0.0);
      } else {
        c = aa / d;
        sv = make_double2(av.x / aa * nu / d, av.y / aa * nu / d);
        r = make_double2(av.x / aa * d, av.y / aa * d);
      }
      Rc[j - 1] = r;
      cs[j - 1] = c;
      sn[j - 1] = sv;
      const double2 gi = gv[j - 1];
      gv[j] = make_double2(-(sv.x * gi.x + sv.y * gi.y), -(sv.x * gi.y - sv.y * gi.x));
      gv[j - 1] = make_double2(c * gi.x, c * gi.y);
      const double est = hypot(gv[j].x, gv[j].y);
      st->res_est = est;
      st->k = j;
      if (a.hist) a.hist[(long)y * a.shist + j] = est;
      const bool bad = !isfinite(est) || !isfinite(nu);
      if (bad) st->status = CGM_ERROR_NONFINITE;
      // the cycle stops at a fifth of the tolerance: in complex64 the true
      // residual of the iterate stays a few 1e-7 above the estimate on the
      // benchmark physics (1.03-1.08e-6 at an estimate of 1e-6, N = 64), so
      // stopping at the tolerance itself would end just above it
      if (bad || est <= 0.2 * st->tol_abs || !(nu > 0.0) || j == a.m) {
        st->done = 1;
        s_done = 1;
      }
    }
    __syncthreads();
  } else if (threadIdx.x == 0 && !(nu > 0.0)) {
    st->done = 1;  // u_0 = 0: nothing to do
    s_done = 1;
  }
  __syncthreads();
  if (s_done || j >= a.m) return;
  // first pass of w against [Q_j, q_j] and the update coefficients
  const double2 hjj = make_double2((sc[1] - str) / nu, (sc[2] - sti) / nu);
  const double2 cc = make_double2(hjj.x / nu, hjj.y / nu);
  for (int i = threadIdx.x; i < j; i += kBS) {
    const double2 si = s[i], ti = t[i];
    h1p[i] = ti;
    a1[i] = si;
    const double2 p = dmul2(si, cc);
    a2[i] = make_double2(ti.x - p.x, ti.y - p.y);
  }
  if (threadIdx.x == 0) {
    h1p[j] = hjj;
    sc[3] = 1.0 / nu;
    sc[4] = cc.x;
    sc[5] = cc.y;
  }
}

// UPD: q_j = (u_j - Q_j a1) / nu and u_{j+1} = w - c u_j - Q_j a2, one
// element per thread, kUU basis loads in flight; ||u_{j+1}||^2 partials,
// the last block forms ||u_{j+1}|| (st->unorm).
__global__ void __launch_bounds__(kBS) k_lr_upd(LrArgs a) {
  DevState* st = boff(a.st, (long)blockIdx.y * a.sst);
  if (st->done) return;
  const int y = blockIdx.y;
  const int j = a.j;
  const long e = (long)blockIdx.x * kBS + threadIdx.x;
  const bool in = e < a.n;
  const float2* Q = a.Q + (long)y * (a.ld + 1) * a.n;
  const double2* a1 = a.a1 + (long)y * a.sSmall;
  const double2* a2 = a.a2 + (long)y * a.sSmall;
  // coefficients and sums in FP64: q_j and u_{j+1} are differences of
  // nearly equal vectors, and a sum of thousands of FP32 terms would carry an
  // error far above the FP32 rounding of the result
  __shared__ double2 c1[kStage / 2], c2[kStage / 2];
  double a1x = 0.0, a1y = 0.0, a2x = 0.0, a2y = 0.0;
  for (int i0 = 0; i0 < j; i0 += kStage / 2) {
    const int cnt = min(kStage / 2, j - i0);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += kBS) {
      c1[i] = a1[i0 + i];
      c2[i] = a2[i0 + i];
    }
    __syncthreads();
    if (in) {
      int i = 0;
      for (; i + kUU <= cnt; i += kUU) {
        float2 q[kUU];
#pragma unroll
        for (int u = 0; u < kUU; ++u) q[u] = nt_load(&Q[(long)(i0 + i + u) * a.n + e]);
#pragma unroll
        for (int u = 0; u < kUU; ++u) {
          const double2 k1 = c1[i + u], k2 = c2[i + u];
          const double qx = q[u].x, qy = q[u].y;
          a1x += k1.x * qx - k1.y * qy;
          a1y += k1.x * qy + k1.y * qx;
          a2x += k2.x * qx - k2.y * qy;
          a2y += k2.x * qy + k2.y * qx;
        }
      }
      for (; i < cnt; ++i) {
        const float2 qv = nt_load(&Q[(long)(i0 + i) * a.n + e]);
        const double2 k1 = c1[i], k2 = c2[i];
        a1x += k1.x * qv.x - k1.y * qv.y;
        a1y += k1.x * qv.y + k1.y * qv.x;
        a2x += k2.x * qv.x - k2.y * qv.y;
        a2y += k2.x * qv.y + k2.y * qv.x;
      }
    }
  }
  const double* sc = a.scal + (long)y * 8;
  double ssq = 0.0;
  if (in) {
    const double inu = sc[3];
    const double ccx = sc[4], ccy = sc[5];
    const float2 uv = a.U[(long)y * a.sU + e];
    double2 wv;
    if (a.W64) {
      wv = a.W64[(long)y * a.n + e];
    } else {
      const float2 w32 = a.W[(long)y * a.sW + e];
      wv = make_double2(w32.x, w32.y);
    }
    float2* qj = a.Q + (long)y * (a.ld + 1) * a.n + (long)j * a.n;
    qj[e] = make_float2((float)((uv.x - a1x) * inu), (float)((uv.y - a1y) * inu));
    const double cux = ccx * uv.x - ccy * uv.y, cuy = ccx * uv.y + ccy * uv.x;
    const float2 un = make_float2((float)(wv.x - cux - a2x), (float)(wv.y - cuy - a2y));
    a.Unext[(long)y * a.sU + e] = un;
    ssq = (double)un.x * un.x + (double)un.y * un.y;
  }
  __shared__ double redn[kBS / 64];
  ssq = wave_sum(ssq);
  if ((threadIdx.x & 63) == 0) redn[threadIdx.x >> 6] = ssq;
  __syncthreads();
  double* partn = a.partn + (long)y * a.sPartn;
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int w = 0; w < kBS / 64; ++w) sum += redn[w];
    put(&partn[blockIdx.x], sum);
  }
  // arrive_last counts against gridDim.x: every block of this system
  if (!arrive_last(&st->tickets[kTicketMisc])) return;
  const double tot = sum_partials(partn, a.nuchunks);
  if (threadIdx.x == 0) {
    st->unorm = sqrt(tot);
    reset_ticket(&st->tickets[kTicketMisc]);
  }
}

// First vector of a cycle: u_0 = r (already in U), ||u_0|| = beta; g = beta e1.
__global__ void k_lr_init(LrArgs a) {
  DevState* st = boff(a.st, (long)blockIdx.y * a.sst);
  if (threadIdx.x == 0) {
    st->unorm = st->beta;
    a.g[(long)blockIdx.y * a.sSmall] = make_double2(st->beta, 0.0);
  }
}

// R y = g[0:k] (R upper triangular, column-major ld), FP64, then the state's
// k is kept in k_last.
__global__ void __launch_bounds__(1024) k_lr_backsub(LrArgs a) {
  DevState* st = boff(a.st, (long)blockIdx.y * a.sst);
  const int y = blockIdx.y;
  const int k = st->k;
  double2* gv = a.g + (long)y * a.sSmall;
  double2* yv = a.y + (long)y * a.sSmall;
  const double2* R = a.R + (long)y * a.ld * a.ld;
  if (threadIdx.x == 0) st->k_last = k;
  __shared__ double2 yi;
  for (int i = k - 1; i >= 0; --i) {
    if (threadIdx.x == 0) {
      const double2 r = R[(long)i * a.ld + i], gi = gv[i];
      const double den = r.x * r.x + r.y * r.y;
      yi = make_double2((gi.x * r.x + gi.y * r.y) / den, (gi.y * r.x - gi.x * r.y) / den);
      yv[i] = yi;
    }
    __syncthreads();
    const double2 yy = yi;
    for (int r = threadIdx.x; r < i; r += blockDim.x) {
      const double2 rv = R[(long)i * a.ld + r];
      double2 gg = gv[r];
      gg.x -= rv.x * yy.x - rv.y * yy.y;
      gg.y -= rv.x * yy.y + rv.y * yy.x;
      gv[r] = gg;
    }
    __syncthreads();
  }
}

// x += Z y over the k_last columns, one element per thread, kUU loads in
// flight.  The sum is accumulated in FP64 with y in FP64: with a strong
// preconditioner the terms y_j z_j can be far larger than their sum, and an
// FP32 sum would leave a rounding error of that larger size in x.
__global__ void __launch_bounds__(kBS) k_lr_xupd(LrArgs a, float2* __restrict__ x, long sx) {
  const DevState* st = boff(a.st, (long)blockIdx.y * a.sst);
  const int y = blockIdx.y;
  const int k = st->k_last;
  const long e = (long)blockIdx.x * kBS + threadIdx.x;
  const float2* Z = a.Z + (long)y * a.ld * a.n;
  const double2* yv = a.y + (long)y * a.sSmall;
  __shared__ double2 cy[kStage / 2];
  double ax = 0.0, ay = 0.0;
  for (int i0 = 0; i0 < k; i0 += kStage / 2) {
    const int cnt = min(kStage / 2, k - i0);
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += kBS) cy[i] = yv[i0 + i];
    __syncthreads();
    if (e < a.n) {
      int i = 0;
      for (; i + kUU <= cnt; i += kUU) {
        float2 z[kUU];
#pragma unroll
        for (int u = 0; u < kUU; ++u) z[u] = Z[(long)(i0 + i + u) * a.n + e];
#pragma unroll
        for (int u = 0; u < kUU; ++u) {
          const double2 c = cy[i + u];
          ax += c.x * z[u].x - c.y * z[u].y;
          ay += c.x * z[u].y + c.y * z[u].x;
        }
      }
      for (; i < cnt; ++i) {
        const float2 zv = Z[(long)(i0 + i) * a.n + e];
        const double2 c = cy[i];
        ax += c.x * zv.x - c.y * zv.y;
        ay += c.x * zv.y + c.y * zv.x;
      }
    }
  }
  if (e < a.n && k > 0) {
    const float2 xv = x[(long)y * sx + e];
    x[(long)y * sx + e] = make_float2((float)(xv.x + ax), (float)(xv.y + ay));
  }
}

// z = u / ||u|| per system (identity preconditioner).
__global__ void __launch_bounds__(kBS) k_lr_scale(const float2* __restrict__ u, long su, float2* __restrict__ z,
                                                  long sz, long n, const DevState* st, long sst) {
  st = boff(st, (long)blockIdx.y * sst);
  if (st->done) return;
  const double d = st->unorm;
  const float r = d > 0.0 ? (float)(1.0 / d) : 0.f;
  const long e = (long)blockIdx.x * kBS + threadIdx.x;
  if (e < n) {
    const float2 v = u[(long)blockIdx.y * su + e];
    z[(long)blockIdx.y * sz + e] = make_float2(v.x * r, v.y * r);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
int lr_alloc(LrWs& w, long n, int m, int B) {
  if (w.m >= m && w.n == n && w.B >= B) return CGM_SUCCESS;
  if (w.m > m && w.n == n) m = w.m;
  if (w.B > B && w.n == n) B = w.B;
  lr_free(w);
  w.n = n;
  w.m = m;
  w.B = B;
  w.nchunks = (int)((n + kBS * kPE - 1) / (kBS * kPE));
  w.nuchunks = (int)((n + kBS - 1) / kBS);
  const int ngroups = (m + kJT - 1) / kJT;
  w.sPart = (long)ngroups * kNV * w.nchunks;
  w.sSmall = m + 1;
  const size_t BB = (size_t)B;
  CGM_HIP(hipMalloc((void**)&w.Q, sizeof(float2) * (size_t)(m + 1) * n * BB));
  CGM_HIP(hipMalloc((void**)&w.Z, sizeof(float2) * (size_t)m * n * BB));
  CGM_HIP(hipMalloc((void**)&w.U, sizeof(float2) * 2 * (size_t)n * BB));
  CGM_HIP(hipMalloc((void**)&w.W, sizeof(float2) * (size_t)n * BB));
  CGM_HIP(hipMalloc((void**)&w.part, sizeof(double2) * (size_t)w.sPart * BB));
  const size_t sm = (size_t)w.sSmall * BB;
  CGM_HIP(hipMalloc((void**)&w.s, sizeof(double2) * sm));
  CGM_HIP(hipMalloc((void**)&w.t, sizeof(double2) * sm));
  CGM_HIP(hipMalloc((void**)&w.h1p, sizeof(double2) * sm));
  CGM_HIP(hipMalloc((void**)&w.g, sizeof(double2) * sm));
  CGM_HIP(hipMalloc((void**)&w.y, sizeof(double2) * sm));
  CGM_HIP(hipMalloc((void**)&w.cs, sizeof(double) * sm));
  CGM_HIP(hipMalloc((void**)&w.sn, sizeof(double2) * sm));
  CGM_HIP(hipMalloc((void**)&w.a1, sizeof(double2) * sm));
  CGM_HIP(hipMalloc((void**)&w.a2, sizeof(double2) * sm));
  CGM_HIP(hipMalloc((void**)&w.scal, sizeof(double) * 8 * BB));
  CGM_HIP(hipMalloc((void**)&w.R, sizeof(double2) * (size_t)m * m * BB));
  CGM_HIP(hipMalloc((void**)&w.partn, sizeof(double) * (size_t)w.nuchunks * BB));
  return CGM_SUCCESS;
}

void lr_free(LrWs& w) {
  void* ptrs[] = {w.Q, w.Z, w.U, w.W, w.part, w.s, w.t, w.h1p, w.g, w.y, w.cs, w.sn, w.a1, w.a2, w.scal, w.R,
                  w.partn};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  w = LrWs();
}

size_t lr_bytes(long n, int m, int B) {
  const long nchunks = (n + kBS * kPE - 1) / (kBS * kPE);
  const long ngroups = (m + kJT - 1) / kJT;
  return (size_t)B * (sizeof(float2) * ((size_t)(2 * m + 4) * n) + sizeof(double2) * (size_t)m * m +
                      sizeof(double2) * (size_t)ngroups * kNV * nchunks + 136 * (size_t)(m + 1));
}

static LrArgs lr_args(const LrWs& w, DevState* st, long sst, double* hist, long shist, int m) {
  LrArgs a{};
  a.m = m;
  a.ld = w.m;
  a.nchunks = w.nchunks;
  a.nuchunks = w.nuchunks;
  a.n = w.n;
  a.Q = w.Q;
  a.Z = w.Z;
  a.sU = 2 * w.n;
  a.sW = w.n;
  a.W = w.W;
  a.W64 = w.W64;
  a.part = w.part;
  a.sPart = w.sPart;
  a.s = w.s;
  a.t = w.t;
  a.h1p = w.h1p;
  a.g = w.g;
  a.y = w.y;
  a.cs = w.cs;
  a.sn = w.sn;
  a.a1 = w.a1;
  a.a2 = w.a2;
  a.scal = w.scal;
  a.sSmall = w.sSmall;
  a.R = w.R;
  a.partn = w.partn;
  a.sPartn = w.nuchunks;
  a.st = st;
  a.sst = sst;
  a.hist = hist;
  a.shist = shist;
  return a;
}

// One long-recurrence cycle of m steps for B systems, from u_0 = r in U[0]
// (its norm in st->beta), then x += Z y.  `precond` computes z_j from
// (u_j, its norm) into Z[j] for all systems; `matvec` computes w = A z_j.
// Every `poll` steps the host reads the states and stops enqueueing once
// every system is done.
int lr_cycle(LrWs& w, hipStream_t st, DevState* dst, DevState* hst, long sst, int B, int m, double* hist,
             long shist, float2* x, long sx, const LrStep& precond, const LrStep& matvec, LrTiming* tm) {
  const long n = w.n;
  LrArgs a = lr_args(w, dst, sst, hist, shist, m);
  const bool vec = (n % 2) == 0;   // 16-byte loads: every vector starts 16-byte aligned
  constexpr int KK = LrTiming::kKernels;
  if (tm && !tm->created) {
    for (int e = 0; e < 2 * KK * LrTiming::kRing; ++e) CGM_HIP(hipEventCreate(&tm->ev[e]));
    tm->created = true;
  }
  int pending = 0;  // ring slots recorded since the last read-back
  auto harvest = [&]() -> int {
    if (!tm || pending == 0) return CGM_SUCCESS;
    const int last = pending - 1;
    // the step's last event: the update's end, or the projection's at the
    // cycle's last step
    CGM_HIP(hipEventSynchronize(tm->ev[2 * KK * last + (tm->slot_has[1][last] ? 3 : 1)]));
    for (int sl = 0; sl < pending; ++sl)
      for (int k = 0; k < KK; ++k) {
        if (!tm->slot_has[k][sl]) continue;
        float t = 0.f;
        CGM_HIP(hipEventElapsedTime(&t, tm->ev[2 * KK * sl + 2 * k], tm->ev[2 * KK * sl + 2 * k + 1]));
        tm->ms[k] += t;
        tm->bytes[k] += tm->slot_bytes[k][sl];
        tm->launches[k] += 1;
      }
    pending = 0;
    return CGM_SUCCESS;
  };
  hipLaunchKernelGGL(k_lr_init, dim3(1, B), dim3(64), 0, st, a);
  CGM_HIP(hipGetLastError());
  const int poll = 32;
  for (int j = 0; j <= m; ++j) {
    a.j = j;
    a.U = w.U + (long)(j & 1) * n;
    a.Unext = w.U + (long)((j + 1) & 1) * n;
    a.ngroups = (j + kJT - 1) / kJT;
    const int with_w = j < m;
    const int sl = pending;
    if (tm)
      for (int k = 0; k < KK; ++k) tm->slot_has[k][sl] = false;
    if (with_w) {
      if (tm) tm->cur = sl;  // the tile application records into this slot
      int rc = precond(j, a.U, 2 * n, w.Z + (long)j * n, (long)w.m * n);
      if (tm) tm->cur = -1;
      if (!rc) rc = matvec(j, w.Z + (long)j * n, (long)w.m * n, w.W, n);
      if (rc) return rc;
    }
    const int gx = (a.ngroups > 0 ? a.ngroups : 1) * w.nchunks;
    if (tm) CGM_HIP(hipEventRecord(tm->ev[2 * KK * sl], st));
    if (vec)
      hipLaunchKernelGGL(k_lr_proj<true>, dim3(gx, B), dim3(kBS), 0, st, a, with_w);
    else
      hipLaunchKernelGGL(k_lr_proj<false>, dim3(gx, B), dim3(kBS), 0, st, a, with_w);
    if (tm) CGM_HIP(hipEventRecord(tm->ev[2 * KK * sl + 1], st));
    hipLaunchKernelGGL(k_lr_red, dim3(a.ngroups > 0 ? a.ngroups : 1, B), dim3(kBS), 0, st, a);
    hipLaunchKernelGGL(k_lr_fin, dim3(1, B), dim3(kBS), 0, st, a);
    if (with_w) {
      if (tm) CGM_HIP(hipEventRecord(tm->ev[2 * KK * sl + 2], st));
      hipLaunchKernelGGL(k_lr_upd, dim3(w.nuchunks, B), dim3(kBS), 0, st, a);
      if (tm) CGM_HIP(hipEventRecord(tm->ev[2 * KK * sl + 3], st));
    }
    CGM_HIP(hipGetLastError());
    if (tm) {
      // algorithmic bytes: PROJ reads q_0..q_{j-1}, u (and w); UPD reads
      // q_0..q_{j-1}, u, w and writes q_j, u_{j+1} (8 B per cell each)
      tm->slot_bytes[0][sl] = 8.0 * (double)n * (j + (with_w ? 2 : 1)) * B;
      tm->slot_bytes[1][sl] = with_w ? 8.0 * (double)n * (j + 4) * B : 0.0;
      tm->slot_has[0][sl] = true;
      tm->slot_has[1][sl] = with_w != 0;
      if (++pending == LrTiming::kRing) {
        const int rc = harvest();
        if (rc) return rc;
      }
    }
    if (j >= 2 * poll && j % poll == 0 && j < m) {
      CGM_HIP(hipMemcpyAsync(hst, dst, sizeof(DevState) * (sst ? B : 1), hipMemcpyDeviceToHost, st));
      CGM_HIP(hipStreamSynchronize(st));
      bool all = true;
      for (int k = 0; k < B && all; ++k) all = boff(hst, (long)k * sst)->done != 0;
      if (all) break;
    }
  }
  if (const int rc = harvest()) return rc;
  hipLaunchKernelGGL(k_lr_backsub, dim3(1, B), dim3(1024), 0, st, a);
  hipLaunchKernelGGL(k_lr_xupd, dim3(w.nuchunks, B), dim3(kBS), 0, st, a, x, sx);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

int lr_scale(hipStream_t st, const float2* u, long su, float2* z, long sz, long n, const DevState* dst, long sst,
             int B) {
  hipLaunchKernelGGL(k_lr_scale, dim3((n + kBS - 1) / kBS, B), dim3(kBS), 0, st, u, su, z, sz, n, dst, sst);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

}  // namespace cgm

// ---------------------------------------------------------------------------
// diagnostics: average device time of the four step kernels at basis size j
// ---------------------------------------------------------------------------
extern "C" int cgm_profile_longrec(cgm_ctx_t c, int j, int reps, float* ms) {
  using namespace cgm;
  CGM_CHECK(c && ms, CGM_ERROR_INVALID_VALUE, "null argument");
  CGM_CHECK(j >= 1 && j < kMaxRestartLR, CGM_ERROR_INVALID_VALUE, "j=%d outside [1, %d)", j, kMaxRestartLR);
  CGM_CHECK(reps >= 1, CGM_ERROR_INVALID_VALUE, "reps < 1");
  CGM_HIP(hipSetDevice(c->device));
  const long n = (long)c->N * c->N;
  int rc = lr_alloc(c->lr, n, j + 1, 1);
  if (rc) return rc;
  LrWs& w = c->lr;
  w.W64 = nullptr;  // the complex64 w (timing only)
  // zero data: the kernels' cost does not depend on the values; the state
  // keeps done = 0 so every launch does its full work (k_lr_fin, which
  // would end the cycle on a zero vector, is timed last)
  CGM_HIP(hipMemsetAsync(w.Q, 0, sizeof(float2) * (size_t)(j + 1) * n, c->stream));
  CGM_HIP(hipMemsetAsync(w.U, 0, sizeof(float2) * 2 * (size_t)n, c->stream));
  CGM_HIP(hipMemsetAsync(w.W, 0, sizeof(float2) * (size_t)n, c->stream));
  CGM_HIP(hipMemsetAsync(w.scal, 0, sizeof(double) * 8, c->stream));
  CGM_HIP(hipMemsetAsync(w.a1, 0, sizeof(double2) * (size_t)w.sSmall, c->stream));
  CGM_HIP(hipMemsetAsync(w.a2, 0, sizeof(double2) * (size_t)w.sSmall, c->stream));
  CGM_HIP(hipMemsetAsync(c->st, 0, sizeof(DevState), c->stream));
  LrArgs a = lr_args(w, c->st, 0, nullptr, 0, j + 1);
  a.j = j;
  a.U = w.U;
  a.Unext = w.U + n;
  a.ngroups = (j + kJT - 1) / kJT;
  hipEvent_t e0, e1;
  CGM_HIP(hipEventCreate(&e0));
  CGM_HIP(hipEventCreate(&e1));
  auto time_it = [&](auto&& launch) -> float {
    launch();
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, c->stream);
    (void)hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    return t / reps;
  };
  const bool vec = (n % 2) == 0;
  ms[0] = time_it([&] {
    if (vec)
      hipLaunchKernelGGL(k_lr_proj<true>, dim3(a.ngroups * w.nchunks, 1), dim3(kBS), 0, c->stream, a, 1);
    else
      hipLaunchKernelGGL(k_lr_proj<false>, dim3(a.ngroups * w.nchunks, 1), dim3(kBS), 0, c->stream, a, 1);
  });
  ms[1] = time_it([&] { hipLaunchKernelGGL(k_lr_red, dim3(a.ngroups, 1), dim3(kBS), 0, c->stream, a); });
  ms[3] = time_it([&] { hipLaunchKernelGGL(k_lr_upd, dim3(w.nuchunks, 1), dim3(kBS), 0, c->stream, a); });
  ms[2] = time_it([&] {
    (void)hipMemsetAsync(&c->st->done, 0, sizeof(int), c->stream);
    hipLaunchKernelGGL(k_lr_fin, dim3(1, 1), dim3(kBS), 0, c->stream, a);
  });
  const hipError_t err = hipGetLastError();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  CGM_HIP(err);
  CGM_HIP(hipStreamSynchronize(c->stream));
  return CGM_SUCCESS;
}

// ---------------------------------------------------------------------------
// live timing of the projection and update kernels inside solves
// ---------------------------------------------------------------------------
extern "C" int cgm_longrec_timing(cgm_ctx_t c, int enable) {
  CGM_CHECK(c, CGM_ERROR_INVALID_VALUE, "null context");
  cgm::LrTiming& t = c->lrt;
  t.on = enable != 0;
  t.cur = -1;
  for (int k = 0; k < cgm::LrTiming::kKernels; ++k) {
    t.ms[k] = t.bytes[k] = 0.0;
    t.launches[k] = 0;
  }
  return CGM_SUCCESS;
}

extern "C" int cgm_longrec_timing_read(cgm_ctx_t c, double* out) {
  CGM_CHECK(c && out, CGM_ERROR_INVALID_VALUE, "null argument");
  const cgm::LrTiming& t = c->lrt;
  for (int k = 0; k < cgm::LrTiming::kKernels; ++k) {
    out[3 * k] = t.ms[k];
    out[3 * k + 1] = t.bytes[k];
    out[3 * k + 2] = (double)t.launches[k];
  }
  return CGM_SUCCESS;
}
