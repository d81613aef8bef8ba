// spmm_util.h - device helpers shared by the CSR and row-block products.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {

constexpr int kBlock = 256;
constexpr int kXcds = 8;

template <int K>
__device__ inline bool all_stopped(const ColState* st) {
  if (st == nullptr) return false;
#pragma unroll
  for (int c = 0; c < K; ++c)
    if (st[c].status == CC_RUNNING) return false;
  return true;
}

// Vector access of CPL consecutive right-hand sides (one 8- or 16-byte load).
template <typename VT, int CPL>
struct VecT;
template <>
struct VecT<double, 1> {
  using T = double;
};
template <>
struct VecT<double, 2> {
  using T = double2;
};
template <>
struct VecT<float, 1> {
  using T = float;
};
template <>
struct VecT<float, 2> {
  using T = float2;
};
template <>
struct VecT<float, 4> {
  using T = float4;
};

template <typename VT, int CPL>
__device__ inline void vload(const VT* __restrict__ p, VT* out) {
  if constexpr (CPL == 1) {
    out[0] = *p;
  } else {
    using V = typename VecT<VT, CPL>::T;
    const V v = *reinterpret_cast<const V*>(p);
    out[0] = v.x;
    out[1] = v.y;
    if constexpr (CPL == 4) {
      out[2] = v.z;
      out[3] = v.w;
    }
  }
}

template <typename VT, int CPL>
__device__ inline void vstore(VT* __restrict__ p, const VT* in) {
  if constexpr (CPL == 1) {
    *p = in[0];
  } else {
    using V = typename VecT<VT, CPL>::T;
    V v;
    v.x = in[0];
    v.y = in[1];
    if constexpr (CPL == 4) {
      v.z = in[2];
      v.w = in[3];
    }
    *reinterpret_cast<V*>(p) = v;
  }
}

// First row r in [0, n] with indptr[r] >= target (indptr is nondecreasing and
// indptr[n] >= target), found by one wave with a 64-ary search: every round
// probes 64 evenly spaced row pointers and keeps the bracket of the first
// one at or past the target. Three rounds cover 262k rows. Every lane of the
// wave returns the same row.
__device__ inline int row_lower_bound(const int* __restrict__ indptr, int n, long long target) {
  const int lane = threadIdx.x & 63;
  int lo = 0, hi = n;
  while (hi - lo > 64) {
    const int step = (hi - lo + 63) / 64;
    const int idx = min(lo + lane * step, hi);
    const unsigned long long m = __ballot(indptr[idx] >= target);
    const int f = m ? __ffsll((long long)m) - 1 : 64;
    const int nlo = f == 0 ? lo : lo + (f - 1) * step;
    const int nhi = f == 64 ? hi : min(lo + f * step, hi);
    lo = nlo;
    hi = nhi;
  }
  const int idx = lo + lane;
  const unsigned long long m = __ballot(idx <= hi && indptr[min(idx, hi)] >= target);
  return m ? lo + __ffsll((long long)m) - 1 : hi;
}

template <bool NTL, typename T>
__device__ inline T stream_ld(const T* p) {
  if constexpr (NTL)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

}  // namespace cc
