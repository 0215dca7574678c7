// Device-resident Arnoldi/GMRES kernels: CGS2 orthogonalisation with
// wavefront-shuffle reductions, fused norm + Givens update in FP64, back
// substitution and the solution update.
//
// Reference counterparts (B/CUDA_GMRES/): the modified Gram-Schmidt loop of
// cublasCdotc + weight_subtract_kernel (GMRES.cuh:289-303, GMRES_kernels.cuh:328-339),
// cublasScnrm2 + get_complex_divided + cublasCsscal (GMRES.cuh:305-321,
// GMRES_kernels.cuh:342-364), the dense-matrix Givens updates with
// cublasCgemm3m (GMRES.cuh:327-371), the LU least-squares solve
// (GMRES.cuh:377-423) and add_kernel (GMRES.cuh:426-434, GMRES_kernels.cuh:547-554).
//
// Reductions: see cgm_reduce.h.
#include <math.h>

#include "cgm_reduce.h"

namespace cgm {

// h[j] = sum_e conj(V_j[e]) w[e], j < nvec.
template <int BS, int CH>
__global__ void __launch_bounds__(BS)
k_dots(const float2* __restrict__ V, long n, int nvec, const float2* __restrict__ w,
       double2* __restrict__ part, int G, double2* __restrict__ out, unsigned* ticket,
       const int* __restrict__ done) {
  if (*done) return;
  for (int c0 = 0; c0 < nvec; c0 += CH) {
    float2 acc[CH];
#pragma unroll
    for (int jj = 0; jj < CH; ++jj) acc[jj] = make_float2(0.f, 0.f);
    for (long e = (long)blockIdx.x * BS + threadIdx.x; e < n; e += (long)G * BS) {
      const float2 wv = w[e];
#pragma unroll
      for (int jj = 0; jj < CH; ++jj) {
        if (c0 + jj < nvec) {
          const float2 v = V[(long)(c0 + jj) * n + e];
          acc[jj].x += v.x * wv.x + v.y * wv.y;
          acc[jj].y += v.x * wv.y - v.y * wv.x;
        }
      }
    }
    block_store<BS, CH>(acc, part, G, c0, nvec);
  }
  if (arrive_last(ticket)) {
    finish_partials<BS>(part, G, nvec, out);
    if (threadIdx.x == 0) reset_ticket(ticket);
  }
}

// w -= sum_j h_j V_j, then h2[j] = sum_e conj(V_j[e]) w[e] from the same
// registers (one read of V for the update and the second projection).
template <int BS, int MAXV>
__global__ void __launch_bounds__(BS)
k_upd_dots(const float2* __restrict__ V, long n, int nvec, float2* __restrict__ w,
           const double2* __restrict__ h, double2* __restrict__ part, int G,
           double2* __restrict__ out, unsigned* ticket, const int* __restrict__ done) {
  if (*done) return;
  float2 hf[MAXV];
#pragma unroll
  for (int j = 0; j < MAXV; ++j) {
    const double2 hv = j < nvec ? h[j] : make_double2(0.0, 0.0);
    hf[j] = make_float2((float)hv.x, (float)hv.y);
  }
  float2 acc[MAXV];
#pragma unroll
  for (int j = 0; j < MAXV; ++j) acc[j] = make_float2(0.f, 0.f);
  for (long e = (long)blockIdx.x * BS + threadIdx.x; e < n; e += (long)G * BS) {
    float2 v[MAXV];
#pragma unroll
    for (int j = 0; j < MAXV; ++j) v[j] = j < nvec ? V[(long)j * n + e] : make_float2(0.f, 0.f);
    float2 wv = w[e];
#pragma unroll
    for (int j = 0; j < MAXV; ++j) {
      wv.x -= hf[j].x * v[j].x - hf[j].y * v[j].y;
      wv.y -= hf[j].x * v[j].y + hf[j].y * v[j].x;
    }
    w[e] = wv;
#pragma unroll
    for (int j = 0; j < MAXV; ++j) {
      acc[j].x += v[j].x * wv.x + v[j].y * wv.y;
      acc[j].y += v[j].x * wv.y - v[j].y * wv.x;
    }
  }
  block_store<BS, MAXV>(acc, part, G, 0, nvec);
  if (arrive_last(ticket)) {
    finish_partials<BS>(part, G, nvec, out);
    if (threadIdx.x == 0) reset_ticket(ticket);
  }
}

// w -= sum_j h_j V_j (any nvec).
template <int BS>
__global__ void __launch_bounds__(BS)
k_upd(const float2* __restrict__ V, long n, int nvec, float2* __restrict__ w,
      const double2* __restrict__ h, int G, const int* __restrict__ done) {
  if (*done) return;
  for (long e = (long)blockIdx.x * BS + threadIdx.x; e < n; e += (long)G * BS) {
    float2 wv = w[e];
    for (int j = 0; j < nvec; ++j) {
      const double2 hv = h[j];
      const float hx = (float)hv.x, hy = (float)hv.y;
      const float2 v = V[(long)j * n + e];
      wv.x -= hx * v.x - hy * v.y;
      wv.y -= hx * v.y + hy * v.x;
    }
    w[e] = wv;
  }
}

// Second CGS pass update w -= sum_j h2_j V_j fused with ||w||^2; the last
// block then forms column i of the Hessenberg matrix (h1 + h2, ||w||),
// applies the previous rotations, builds rotation i and updates g, all in
// FP64.  The rotation is the unitary [[c, s], [-conj(s), c]] with real c,
// chosen so that it annihilates H[i+1, i].
template <int BS>
__global__ void __launch_bounds__(BS)
k_upd_norm_givens(const float2* __restrict__ V, long n, int i, float2* __restrict__ w,
                  KrylovWs ws, DevState* st, double* hist, int m, const int* __restrict__ done) {
  if (*done) return;
  const int nvec = i + 1;
  double ss = 0.0;
  for (long e = (long)blockIdx.x * BS + threadIdx.x; e < n; e += (long)ws.G * BS) {
    float2 wv = w[e];
    for (int j = 0; j < nvec; ++j) {
      const double2 hv = ws.h2[j];
      const float hx = (float)hv.x, hy = (float)hv.y;
      const float2 v = V[(long)j * n + e];
      wv.x -= hx * v.x - hy * v.y;
      wv.y -= hx * v.y + hy * v.x;
    }
    w[e] = wv;
    ss += (double)wv.x * wv.x + (double)wv.y * wv.y;
  }
  __shared__ double red[BS / 64];
  ss = wave_sum(ss);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int k = 0; k < BS / 64; ++k) s += red[k];
    put(&ws.partn[blockIdx.x], s);
  }
  if (!arrive_last(&st->tickets[kTicketNorm])) return;

  __shared__ double2 col[kMaxRestart + 1];
  __shared__ double2 lsn[kMaxRestart];
  __shared__ double lcs[kMaxRestart];
  __shared__ double s_norm;
  {
    const double sq = sum_partials(ws.partn, ws.G);
    if (threadIdx.x == 0) s_norm = sqrt(sq);
  }
  for (int j = threadIdx.x; j <= i; j += BS) {
    const double2 a = ws.h1[j], b = ws.h2[j];
    col[j] = make_double2(a.x + b.x, a.y + b.y);
    if (j < i) {
      lcs[j] = ws.cs[j];
      lsn[j] = ws.sn[j];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double hn = s_norm;
    col[i + 1] = make_double2(hn, 0.0);
    for (int j = 0; j < i; ++j) {
      const double c = lcs[j];
      const double2 s = lsn[j];
      const double2 a = col[j], b = col[j + 1];
      // [a; b] <- [c a + s b; -conj(s) a + c b]
      col[j] = make_double2(c * a.x + (s.x * b.x - s.y * b.y), c * a.y + (s.x * b.y + s.y * b.x));
      col[j + 1] = make_double2(-(s.x * a.x + s.y * a.y) + c * b.x, -(s.x * a.y - s.y * a.x) + c * b.y);
    }
    const double2 a = col[i];
    const double aa = hypot(a.x, a.y);
    const double d = hypot(aa, hn);
    double c;
    double2 s, r;
    if (aa == 0.0) {
      c = 0.0;
      s = make_double2(1.0, 0.0);
      r = make_double2(hn, 0.0);
    } else {
      c = aa / d;
      // s = (a/|a|) conj(hn) / d, hn real
      s = make_double2(a.x / aa * hn / d, a.y / aa * hn / d);
      r = make_double2(a.x / aa * d, a.y / aa * d);
    }
    col[i] = r;
    ws.cs[i] = c;
    ws.sn[i] = s;
    const double2 gi = ws.g[i];
    // g[i+1] = -conj(s) g[i], g[i] = c g[i]
    ws.g[i + 1] = make_double2(-(s.x * gi.x + s.y * gi.y), -(s.x * gi.y - s.y * gi.x));
    ws.g[i] = make_double2(c * gi.x, c * gi.y);
    const double est = hypot(ws.g[i + 1].x, ws.g[i + 1].y);
    st->hn = hn;
    st->k = i + 1;
    st->res_est = est;
    if (hist) hist[i] = est;
    const bool bad = !isfinite(hn) || !isfinite(est);
    if (bad) st->status = CGM_ERROR_NONFINITE;
    if (bad || !(hn > 0.0) || est <= st->tol_abs) st->done = 1;
  }
  __syncthreads();
  for (int j = threadIdx.x; j <= i; j += BS) ws.R[(long)j * m + i] = col[j];
  if (threadIdx.x == 0) reset_ticket(&st->tickets[kTicketNorm]);
}

// ||r|| of a complex64 vector.  what = 0: start of a cycle (beta = ||r||,
// g = beta e1, done/k reset); what = 1: ||b|| and the absolute tolerance;
// what = 2: result to norm_out only.
__global__ void __launch_bounds__(kKB)
k_norm_init(const float2* __restrict__ r, long n, double* partn, int G, DevState* st,
            double2* g, int what, double tol, long sr, long spartn, long ssmall, long sst) {
  {
    const long k = blockIdx.y;
    r += k * sr;
    partn += k * spartn;
    g += k * ssmall;
    st = boff(st, k * sst);
  }
  double ss = 0.0;
  for (long e = (long)blockIdx.x * kKB + threadIdx.x; e < n; e += (long)G * kKB) {
    const float2 v = r[e];
    ss += (double)v.x * v.x + (double)v.y * v.y;
  }
  __shared__ double red[kKB / 64];
  ss = wave_sum(ss);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int k = 0; k < kKB / 64; ++k) s += red[k];
    put(&partn[blockIdx.x], s);
  }
  if (!arrive_last(&st->tickets[kTicketInit])) return;
  {
    const double s = sum_partials(partn, G);
    if (threadIdx.x == 0) {
      const double nrm = sqrt(s);
      if (what == 1) {
        st->bnorm = nrm;
        st->tol_abs = tol * nrm;
        st->status = isfinite(nrm) ? 0 : CGM_ERROR_NONFINITE;
      } else if (what == 0) {
        st->beta = nrm;
        g[0] = make_double2(nrm, 0.0);
        st->k = 0;
        st->res_est = nrm;
        const bool bad = !isfinite(nrm);
        if (bad) st->status = CGM_ERROR_NONFINITE;
        st->done = (bad || nrm <= st->tol_abs || !(nrm > 0.0)) ? 1 : 0;
      } else {
        st->norm_out = nrm;
      }
      reset_ticket(&st->tickets[kTicketInit]);
    }
  }
}

// ||v|| of a complex128 vector -> st->norm_out.
__global__ void __launch_bounds__(kKB)
k_norm64(const double2* __restrict__ v, long n, double* partn, int G, DevState* st) {
  double ss = 0.0;
  for (long e = (long)blockIdx.x * kKB + threadIdx.x; e < n; e += (long)G * kKB) {
    const double2 a = v[e];
    ss += a.x * a.x + a.y * a.y;
  }
  __shared__ double red[kKB / 64];
  ss = wave_sum(ss);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ss;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int k = 0; k < kKB / 64; ++k) s += red[k];
    put(&partn[blockIdx.x], s);
  }
  if (!arrive_last(&st->tickets[kTicketMisc])) return;
  {
    const double s = sum_partials(partn, G);
    if (threadIdx.x == 0) {
      st->norm_out = sqrt(s);
      reset_ticket(&st->tickets[kTicketMisc]);
    }
  }
}

// Solve R y = g[0:k] (upper triangular, row-major ld m) in FP64.
__global__ void __launch_bounds__(kKB) k_backsub(KrylovWs ws, DevState* st, int m, long sst) {
  {
    const long b = blockIdx.y;
    ws.g += b * ws.sSmall;
    ws.R += b * ws.sSmall;
    ws.y += b * ws.sSmall;
    st = boff(st, b * sst);
  }
  const int k = st->k;
  if (threadIdx.x == 0) st->k_last = k;
  __shared__ double2 gs[kMaxRestart];
  __shared__ double2 yi;
  for (int j = threadIdx.x; j < k; j += kKB) gs[j] = ws.g[j];
  __syncthreads();
  for (int i = k - 1; i >= 0; --i) {
    if (threadIdx.x == 0) {
      const double2 r = ws.R[(long)i * m + i], gi = gs[i];
      const double den = r.x * r.x + r.y * r.y;
      yi = make_double2((gi.x * r.x + gi.y * r.y) / den, (gi.y * r.x - gi.x * r.y) / den);
      ws.y[i] = yi;
    }
    __syncthreads();
    const double2 yv = yi;
    for (int j = threadIdx.x; j < i; j += kKB) {
      const double2 r = ws.R[(long)j * m + i];
      gs[j].x -= r.x * yv.x - r.y * yv.y;
      gs[j].y -= r.x * yv.y + r.y * yv.x;
    }
    __syncthreads();
  }
}

// x += sum_{j < k} y_j V_j
__global__ void __launch_bounds__(kKB)
k_xupd(const float2* __restrict__ V, long n, float2* __restrict__ x, const double2* __restrict__ y,
       const DevState* st, int G, long sV, long sx, long sy, long sst) {
  {
    const long b = blockIdx.y;
    V += b * sV;
    x += b * sx;
    y += b * sy;
    st = boff(st, b * sst);
  }
  const int k = st->k;
  for (long e = (long)blockIdx.x * kKB + threadIdx.x; e < n; e += (long)G * kKB) {
    float2 xv = x[e];
    for (int j = 0; j < k; ++j) {
      const double2 yv = y[j];
      const float yx = (float)yv.x, yy = (float)yv.y;
      const float2 v = V[(long)j * n + e];
      xv.x += yx * v.x - yy * v.y;
      xv.y += yx * v.y + yy * v.x;
    }
    x[e] = xv;
  }
}

// ---------------------------------------------------------------------------
int krylov_alloc(KrylovWs& ws, long n, int m, hipStream_t st, int B) {
  if (ws.m >= m && ws.n == n && ws.B >= B) return CGM_SUCCESS;
  if (ws.m > m) m = ws.m;
  if (ws.B > B) B = ws.B;
  krylov_free(ws);
  ws.n = n;
  ws.m = m;
  ws.B = B;
  long blocks = (n + kKB - 1) / kKB;
  ws.G = (int)(blocks < 256 ? blocks : 256);
  if (ws.G < 1) ws.G = 1;
  const size_t mm = (size_t)m;
  const size_t BB = (size_t)B;
  int gmax = ws.G;
  if (fused_supported((int)lrint(sqrt((double)n)), 1)) {
    const int gf = fused_blocks((int)lrint(sqrt((double)n)));
    if (gf > gmax) gmax = gf;
  }
  // per-system strides; every small FP64 array gets the largest small size
  ws.sBasis = (long)n * (long)(mm + 1);
  ws.sSmall = (long)((mm + 1) * mm > 4 ? (mm + 1) * mm : 4);
  ws.sPart = (long)(2 * mm + 4) * gmax;
  ws.sPartn = gmax;
  ws.sVec = n;
  const size_t sm = (size_t)ws.sSmall;
  CGM_HIP(hipMalloc((void**)&ws.V, sizeof(float2) * (size_t)ws.sBasis * BB));
  CGM_HIP(hipMalloc((void**)&ws.h1, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.h2, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.cs, sizeof(double) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.sn, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.g, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.R, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.y, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.part, sizeof(double2) * (size_t)ws.sPart * BB));
  CGM_HIP(hipMalloc((void**)&ws.partn, sizeof(double) * (size_t)ws.sPartn * BB));
  CGM_HIP(hipMalloc((void**)&ws.H, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.h1p, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.a1, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.a2, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.scal, sizeof(double2) * sm * BB));
  CGM_HIP(hipMalloc((void**)&ws.U, sizeof(float2) * 2 * (size_t)n * BB));
  CGM_HIP(hipMalloc((void**)&ws.Z, sizeof(float2) * (size_t)n * BB));
  (void)st;
  return CGM_SUCCESS;
}

void krylov_free(KrylovWs& ws) {
  void* ptrs[] = {ws.V, ws.h1, ws.h2, ws.cs, ws.sn, ws.g, ws.R, ws.y, ws.part, ws.partn,
                  ws.H, ws.h1p, ws.a1, ws.a2, ws.scal, ws.U, ws.Z};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  ws = KrylovWs();
}

int launch_norm_init(hipStream_t st, const KrylovWs& ws, const float2* r, DevState* s, int what,
                     double tol, long sr, long sst, int nsys) {
  const int B = nsys;
  hipLaunchKernelGGL(k_norm_init, dim3(ws.G, B), dim3(kKB), 0, st, r, ws.n, ws.partn, ws.G, s, ws.g,
                     what, tol, sr, ws.sPartn, ws.sSmall, sst);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

int launch_norm64(hipStream_t st, const double2* v, long n, double* partn, int G, DevState* s) {
  hipLaunchKernelGGL(k_norm64, dim3(G), dim3(kKB), 0, st, v, n, partn, G, s);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

// CGS2 for step i: w = V_{i+1} against V_0..V_i, then ||w|| and rotation i.
// part: 0 first projection, 1 update + second projection, 2 update + norm +
// Givens, -1 all three.
int launch_cgs2_step(hipStream_t st, const KrylovWs& ws, int i, DevState* s, double* hist,
                     int part) {
  const int nvec = i + 1;
  const long n = ws.n;
  float2* w = ws.V + (long)(i + 1) * n;
  const int* done = &s->done;
  if (part < 0 || part == 0)
    hipLaunchKernelGGL((k_dots<kKB, 16>), dim3(ws.G), dim3(kKB), 0, st, ws.V, n, nvec, w, ws.part,
                       ws.G, ws.h1, &s->tickets[kTicketDots1], done);
  if (part < 0 || part == 1) {
    if (nvec <= 16) {
      hipLaunchKernelGGL((k_upd_dots<kKB, 16>), dim3(ws.G), dim3(kKB), 0, st, ws.V, n, nvec, w,
                         ws.h1, ws.part, ws.G, ws.h2, &s->tickets[kTicketDots2], done);
    } else if (nvec <= 32) {
      hipLaunchKernelGGL((k_upd_dots<kKB / 2, 32>), dim3(ws.G), dim3(kKB / 2), 0, st, ws.V, n, nvec,
                         w, ws.h1, ws.part, ws.G, ws.h2, &s->tickets[kTicketDots2], done);
    } else {
      hipLaunchKernelGGL((k_upd<kKB>), dim3(ws.G), dim3(kKB), 0, st, ws.V, n, nvec, w, ws.h1, ws.G,
                         done);
      hipLaunchKernelGGL((k_dots<kKB, 16>), dim3(ws.G), dim3(kKB), 0, st, ws.V, n, nvec, w,
                         ws.part, ws.G, ws.h2, &s->tickets[kTicketDots2], done);
    }
  }
  if (part < 0 || part == 2)
    hipLaunchKernelGGL((k_upd_norm_givens<kKB>), dim3(ws.G), dim3(kKB), 0, st, ws.V, n, i, w, ws,
                       s, hist, ws.m, done);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

int launch_backsub_update(hipStream_t st, const KrylovWs& ws, DevState* s, float2* x, const float2* basis,
                          long sx, long sst, int nsys) {
  const int B = nsys;
  hipLaunchKernelGGL(k_backsub, dim3(1, B), dim3(kKB), 0, st, ws, s, ws.m, sst);
  hipLaunchKernelGGL(k_xupd, dim3(ws.G, B), dim3(kKB), 0, st, basis ? basis : ws.V, ws.n, x, ws.y, s, ws.G,
                     ws.sBasis, sx, ws.sSmall, sst);
  CGM_HIP(hipGetLastError());
  return CGM_SUCCESS;
}

}  // namespace cgm
