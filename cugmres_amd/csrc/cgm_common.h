// Shared device/host helpers for the gfx950 GMRES library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/cugmres.h"

#define CGM_HIP(call)                                                         \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      cgm_set_last_error("HIP error %s at %s:%d: %s", hipGetErrorName(e_),    \
                         __FILE__, __LINE__, hipGetErrorString(e_));          \
      return CGM_ERROR_DEVICE;                                                \
    }                                                                         \
  } while (0)

#define CGM_CHECK(cond, code, ...)        \
  do {                                    \
    if (!(cond)) {                        \
      cgm_set_last_error(__VA_ARGS__);    \
      return (code);                      \
    }                                     \
  } while (0)

void cgm_set_last_error(const char* fmt, ...);

namespace cgm {

// Complex helpers on float2/double2 (interleaved re, im; layout of
// cgm_complex64/cgm_complex128 and of the reference's cuComplex).
template <typename T> struct c2;
template <> struct c2<float> { using type = float2; };
template <> struct c2<double> { using type = double2; };
template <typename T> using cplx = typename c2<T>::type;

__host__ __device__ inline float2 mk(float a, float b) { return make_float2(a, b); }
__host__ __device__ inline double2 mk(double a, double b) { return make_double2(a, b); }
template <typename V> __host__ __device__ inline V cadd(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V> __host__ __device__ inline V csub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename V> __host__ __device__ inline V cmul(V a, V b) {
  V r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
template <typename V, typename S> __host__ __device__ inline V cscale(V a, S s) { V r; r.x = a.x * s; r.y = a.y * s; return r; }

// Complex elements one M-point transform occupies in LDS: single-precision
// transforms keep one pad element after every 8 (see cgm_fft.h), double
// transforms are unpadded.
inline int lds_len(int M, bool fp64) { return fp64 ? M : M + M / 8; }

// Work-group id remapped so that consecutive logical tiles land on the same
// XCD: the dispatcher places work-group b on XCD (b % 8), so logical tile
// (b % 8) * (G / 8) + b / 8 gives each XCD a contiguous range of tiles and
// therefore shared L2 lines.  Falls back to identity when G % 8 != 0.
__device__ inline int xcd_tile(int b, int G) {
  if ((G & 7) != 0) return b;
  return (b & 7) * (G >> 3) + (b >> 3);
}

}  // namespace cgm
