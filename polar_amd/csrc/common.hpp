// Shared helpers for the gfx950 kernels: error state, vector loads, wave/group geometry.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>

#include <algorithm>
#include <initializer_list>
#include <utility>

#include "../../include/polar_amd.h"

namespace polar {

int set_error(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, int line);
// Clears `bytes` bytes (a multiple of 4) at device address p with a fill kernel on stream st;
// used instead of hipMemsetAsync, so a caller that captures the stream records no memset node.
int zero_async(void* p, size_t bytes, hipStream_t st);

#define POLAR_HIP(x)                                   \
  do {                                                 \
    hipError_t e_ = (x);                               \
    if (e_ != hipSuccess) return hip_fail(e_, #x, __LINE__); \
  } while (0)

#define POLAR_LAUNCH_CHECK() POLAR_HIP(hipGetLastError())

constexpr int kBlock = 256;        // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kMaxGrid = 1024;     // fixed grid cap: the deterministic reductions sum <= 1024 partials

inline int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// Vector type of CPL consecutive elements (16-byte loads for double2 / float4).
template <typename T, int CPL> struct VecT;
template <> struct VecT<double, 1> { using type = double; };
template <> struct VecT<double, 2> { using type = double2; };
template <> struct VecT<double, 4> { using type = double4; };   // FP64 rows beside FP32 vectors
template <> struct VecT<float, 1> { using type = float; };
template <> struct VecT<float, 2> { using type = float2; };
template <> struct VecT<float, 4> { using type = float4; };

template <typename T, int CPL>
__device__ __forceinline__ void vload(const T* __restrict__ p, T (&o)[CPL]) {
  using V = typename VecT<T, CPL>::type;
  V v = *reinterpret_cast<const V*>(p);
  const T* e = reinterpret_cast<const T*>(&v);
#pragma unroll
  for (int i = 0; i < CPL; ++i) o[i] = e[i];
}

template <typename T, int CPL>
__device__ __forceinline__ void vstore(T* __restrict__ p, const T (&o)[CPL]) {
  using V = typename VecT<T, CPL>::type;
  V v;
  T* e = reinterpret_cast<T*>(&v);
#pragma unroll
  for (int i = 0; i < CPL; ++i) e[i] = o[i];
  *reinterpret_cast<V*>(p) = v;
}

// Non-temporal forms, for data that one kernel touches exactly once (kept out of L2).
template <typename T, int CPL>
__device__ __forceinline__ void vload_nt(const T* __restrict__ p, T (&o)[CPL]) {
  if constexpr (CPL == 1) {
    o[0] = __builtin_nontemporal_load(p);
  } else {
    typedef T V __attribute__((ext_vector_type(CPL)));
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
#pragma unroll
    for (int i = 0; i < CPL; ++i) o[i] = v[i];
  }
}

template <typename T, int CPL>
__device__ __forceinline__ void vstore_nt(T* __restrict__ p, const T (&o)[CPL]) {
  if constexpr (CPL == 1) {
    __builtin_nontemporal_store(o[0], p);
  } else {
    typedef T V __attribute__((ext_vector_type(CPL)));
    V v;
#pragma unroll
    for (int i = 0; i < CPL; ++i) v[i] = o[i];
    __builtin_nontemporal_store(v, reinterpret_cast<V*>(p));
  }
}

// Row-lane geometry: L lanes per row (power of two <= 64), each lane owns CPL consecutive
// columns c0 = li*CPL; G = 64/L row-groups per wave.
struct Geo {
  int L, G, li, gi, wave, lane;
  __device__ __forceinline__ Geo(int L_) : L(L_) {
    lane = threadIdx.x & 63;
    wave = threadIdx.x >> 6;
    G = 64 / L;
    li = lane & (L - 1);
    gi = lane / L;
  }
};

// Host: choose lanes per row for R columns at CPL columns per lane.
inline int lanes_for(int64_t R, int cpl) {
  int64_t need = (R + cpl - 1) / cpl;
  int L = 1;
  while (L < need) L <<= 1;
  return L;
}

template <typename T> struct DT;
template <> struct DT<double> { static constexpr int code = POLAR_F64; };
template <> struct DT<float> { static constexpr int code = POLAR_F32; };

// Double-double accumulation (error-free TwoSum; the error terms are carried in `lo`).
struct DD {
  double hi, lo;
};

__device__ __forceinline__ void dd_add(DD& a, double b) {
#pragma clang fp contract(off)
  const double s = a.hi + b;
  const double bb = s - a.hi;
  const double err = (a.hi - (s - bb)) + (b - bb);
  a.hi = s;
  a.lo += err;
}

__device__ __forceinline__ DD dd_merge(DD a, DD b) {
#pragma clang fp contract(off)
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b.hi - bb);
  e += a.lo + b.lo;
  DD r;
  r.hi = s + e;
  r.lo = e - (r.hi - s);
  return r;
}

// Merge a double-double across the L lanes of a row-group (L a power of two); all lanes get it.
__device__ __forceinline__ DD dd_group_sum(DD v, int L) {
  for (int o = L >> 1; o > 0; o >>= 1) {
    DD w;
    w.hi = __shfl_xor(v.hi, o, 64);
    w.lo = __shfl_xor(v.lo, o, 64);
    v = dd_merge(v, w);
  }
  return v;
}

inline int elem_size(int dtype) { return dtype == POLAR_F64 ? 8 : 4; }
inline bool dtype_ok(int dtype) { return dtype == POLAR_F64 || dtype == POLAR_F32; }

}  // namespace polar
