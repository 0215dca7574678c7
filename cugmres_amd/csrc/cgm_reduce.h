// Deterministic cross-block reductions shared by the Krylov kernels.
//
// Every block writes its partial sums, the last block to finish (ticket
// counter) combines them in a fixed order and resets the counter, so repeated
// runs are bitwise identical and nothing has to be cleared between launches.
// The partials are the only data that crosses blocks inside a kernel; they are
// written and read with agent-scope atomic stores/loads (coherent across the
// XCDs' L2 caches) and the writer waits for its stores before taking a ticket.
// A device-wide fence would instead write back and invalidate the whole L2
// in every block.
#pragma once
#include "cgm_internal.h"

namespace cgm {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of a float over the 64 lanes of a wave with DPP moves (VALU, no LDS
// traffic): rotate-by-8 and rotate-by-4 inside each row of 16 lanes, xor-2 and
// xor-1 quad permutes, then the four row sums are read out with readlane.
// Every lane gets the same value; the summation order is fixed.
template <int CTRL>
__device__ inline float dpp_mov(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
__device__ inline float wave_sum_f(float v) {
  v += dpp_mov<0x128>(v);  // row_ror:8
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
  const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
  const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
  const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
  return (r0 + r1) + (r2 + r3);
}

// Sum over each half (32 lanes) of a wave: after the row-local DPP steps
// every lane of a 16-lane row holds the row sum; rows 0+1 form the first
// half, rows 2+3 the second.  Lanes of a half all get the half's sum.
__device__ inline float half_wave_sum_f(float v) {
  v += dpp_mov<0x128>(v);  // row_ror:8
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
  const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
  const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
  const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
  return (threadIdx.x & 32) ? (r2 + r3) : (r0 + r1);
}

__device__ inline void put(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double get(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void put2(double2* p, double re, double im) {
  put(&p->x, re);
  put(&p->y, im);
}
__device__ inline double2 get2(const double2* p) { return make_double2(get(&p->x), get(&p->y)); }

// True in the last block to arrive.  Every thread first waits for its own
// outstanding memory operations (the partial stores above), so the partials
// of all blocks are complete when the last ticket is drawn.
__device__ inline bool arrive_last(unsigned* ticket) {
  __shared__ int s_last;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (t == gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  return s_last != 0;
}

__device__ inline void reset_ticket(unsigned* ticket) {
  __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduce CH complex float accumulators across the block and store them as
// part[(c0 + jj) * G + blockIdx.x] (FP64).
template <int BS, int CH>
__device__ inline void block_store(const float2 (&acc)[CH], double2* part, int G, int c0, int nvec) {
  __shared__ double red[BS / 64][CH][2];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int jj = 0; jj < CH; ++jj) {
    if (c0 + jj < nvec) {
      const float re = wave_sum_f(acc[jj].x);
      const float im = wave_sum_f(acc[jj].y);
      if (lane == 0) { red[wave][jj][0] = re; red[wave][jj][1] = im; }
    }
  }
  __syncthreads();
  if (threadIdx.x < CH && c0 + (int)threadIdx.x < nvec) {
    double re = 0, im = 0;
#pragma unroll
    for (int w = 0; w < BS / 64; ++w) { re += red[w][threadIdx.x][0]; im += red[w][threadIdx.x][1]; }
    put2(&part[(long)(c0 + threadIdx.x) * G + blockIdx.x], re, im);
  }
  __syncthreads();
}

// out[j] = sum_g part[j * G + g] in a fixed order (last block only).  Each
// j gets a group of L lanes (the largest power of two <= 64 with
// L * nvec <= BS); a lane
// issues all of its loads before summing, then the group combines its lanes
// with butterfly shuffles.
template <int BS>
__device__ inline void finish_partials(const double2* part, int G, int nvec, double2* out) {
  int np2 = 1;
  while (np2 < nvec) np2 <<= 1;
  int L = BS / np2;
  L = L > 64 ? 64 : (L < 1 ? 1 : L);
  const int lanes_total = BS / L;  // j values handled per round
  const int sub = threadIdx.x % L;
  for (int j0 = 0; j0 < nvec; j0 += lanes_total) {
    const int j = j0 + threadIdx.x / L;
    double re = 0, im = 0;
    if (j < nvec) {
      constexpr int U = 8;
      for (int g0 = sub; g0 < G; g0 += U * L) {
        double2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int g = g0 + u * L;
          v[u] = g < G ? get2(&part[(long)j * G + g]) : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) { re += v[u].x; im += v[u].y; }
      }
    }
    for (int off = 1; off < L; off <<= 1) {
      re += __shfl_xor(re, off, 64);
      im += __shfl_xor(im, off, 64);
    }
    if (sub == 0 && j < nvec) out[j] = make_double2(re, im);
  }
}

// sum of G doubles (agent-scope partials) by the first wave, fixed order.
__device__ inline double sum_partials(const double* partn, int G) {
  double s = 0.0;
  if (threadIdx.x < 64) {
    constexpr int U = 8;
    for (int g0 = threadIdx.x; g0 < G; g0 += U * 64) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = (g0 + u * 64 < G) ? get(&partn[g0 + u * 64]) : 0.0;
#pragma unroll
      for (int u = 0; u < U; ++u) s += v[u];
    }
    s = wave_sum(s);
  }
  return s;
}

}  // namespace cgm
