// rowblock.hip - row-block form of a CSR matrix and its product with k
// right-hand sides.
//
// With k = 16 right-hand sides a CSR product gathers one 64-byte (float32)
// or 128-byte (float64) row of X per nonzero, and those gathers, not the
// matrix stream, bound it: on the 3D B-spline systems a row has ~340
// nonzeros and neighbouring rows share most of their columns. The row-block
// form takes the rows in groups of CC_RB_ROWS = 4 consecutive rows and
// stores, per group, the ascending union of their columns and, per union
// column, the 4 row values (zero where a row has no entry). One gathered X
// row then serves 4 rows: on the 81k-unknown p=3 system the union holds 0.36
// of the nonzeros, so the gathers drop 2.8x while the stored bytes (values
// padded with zeros, one index per 4 values) stay within 10% of CSR's.
//
// The form is built on the device from a CSR matrix with ascending column
// indices in each row (duplicates are summed): a count pass (one thread per
// group merges its 4 sorted rows), a one-block scan, and a fill pass.

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "../../include/cutcell_hip.h"
#include "common.h"
#include "spmm_util.h"

namespace cc {

constexpr int kRb = CC_RB_ROWS;

// Merge the sorted rows of group g. FILL = false: write the union size to
// count[g] (-1 if a row's indices descend). FILL = true: write the union
// columns and the per-row values at gptr[g].
template <bool FILL>
__global__ __launch_bounds__(256) void k_rb_merge(int n, int ng, const int* __restrict__ indptr,
                                                  const int* __restrict__ indices, const double* __restrict__ vals,
                                                  int* __restrict__ gptr, int* __restrict__ gcol,
                                                  double* __restrict__ gval, float* __restrict__ gval32) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  int p[kRb], e[kRb];
#pragma unroll
  for (int r = 0; r < kRb; ++r) {
    const int row = g * kRb + r;
    p[r] = row < n ? indptr[row] : 0;
    e[r] = row < n ? indptr[row + 1] : 0;
  }
  int out = FILL ? gptr[g] : 0;
  bool bad = false;
  while (true) {
    int m = INT_MAX;
#pragma unroll
    for (int r = 0; r < kRb; ++r)
      if (p[r] < e[r]) m = min(m, indices[p[r]]);
    if (m == INT_MAX) break;
    double v[kRb];
#pragma unroll
    for (int r = 0; r < kRb; ++r) {
      v[r] = 0.0;
      while (p[r] < e[r] && indices[p[r]] == m) {
        if (FILL) v[r] += vals[p[r]];
        ++p[r];
      }
      // Every adjacent pair of a row is seen here once: its first entry was
      // the smallest head, so the next one must not be below it.
      if (p[r] < e[r] && indices[p[r]] < m) bad = true;
    }
    if (FILL) {
      gcol[out] = m;
#pragma unroll
      for (int r = 0; r < kRb; ++r) {
        if (gval) gval[(size_t)out * kRb + r] = v[r];
        if (gval32) gval32[(size_t)out * kRb + r] = (float)v[r];
      }
    }
    ++out;
  }
  if (!FILL) gptr[g] = bad ? -1 : out;
}

// ---------------------------------------------------------------------------
// Paired row blocks: 8-row groups made of two 4-row groups (pairs[2 p] and
// pairs[2 p + 1]; the second -1 for a group left alone), the ascending union
// of their 8 rows' columns and 8 values per union column. On the 3D B-spline
// systems a group's best partner is the group of the same x-range a few grid
// lines away in y: on the 81k planar system the union then holds 0.23-0.25 of
// the nonzeros against 0.36 for 4-row groups, so the gathered X rows drop by
// a third while the stored values grow by about an eighth.
// ---------------------------------------------------------------------------
constexpr int kRb2 = 2 * kRb;

// Row i (0..7) of paired group p, or -1.
__device__ inline int pair_row(const int* __restrict__ pairs, int p, int i, int n) {
  const int g = pairs[2 * p + (i >= kRb ? 1 : 0)];
  const int row = g < 0 ? -1 : g * kRb + (i % kRb);
  return row < n ? row : -1;
}

template <bool FILL>
__global__ __launch_bounds__(256) void k_rb8_merge(int n, int np, const int* __restrict__ pairs,
                                                   const int* __restrict__ indptr, const int* __restrict__ indices,
                                                   const double* __restrict__ vals, int* __restrict__ gptr,
                                                   int* __restrict__ gcol, double* __restrict__ gval,
                                                   float* __restrict__ gval32) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= np) return;
  int p[kRb2], e[kRb2];
#pragma unroll
  for (int r = 0; r < kRb2; ++r) {
    const int row = pair_row(pairs, g, r, n);
    p[r] = row >= 0 ? indptr[row] : 0;
    e[r] = row >= 0 ? indptr[row + 1] : 0;
  }
  int out = FILL ? gptr[g] : 0;
  bool bad = false;
  while (true) {
    int m = INT_MAX;
#pragma unroll
    for (int r = 0; r < kRb2; ++r)
      if (p[r] < e[r]) m = min(m, indices[p[r]]);
    if (m == INT_MAX) break;
    double v[kRb2];
#pragma unroll
    for (int r = 0; r < kRb2; ++r) {
      v[r] = 0.0;
      while (p[r] < e[r] && indices[p[r]] == m) {
        if (FILL) v[r] += vals[p[r]];
        ++p[r];
      }
      if (p[r] < e[r] && indices[p[r]] < m) bad = true;
    }
    if (FILL) {
      gcol[out] = m;
#pragma unroll
      for (int r = 0; r < kRb2; ++r) {
        if (gval) gval[(size_t)out * kRb2 + r] = v[r];
        if (gval32) gval32[(size_t)out * kRb2 + r] = (float)v[r];
      }
    }
    ++out;
  }
  if (!FILL) gptr[g] = bad ? -1 : out;
}

// One block of 1024 threads: counts gptr[0, ng) -> exclusive offsets
// gptr[0, ng], gptr[ng] = total, or -1 if a count was negative or the total
// does not fit an int32.
__global__ __launch_bounds__(1024) void k_rb_scan(int ng, int* __restrict__ gptr) {
  __shared__ long long sums[1024];
  __shared__ int bad;
  const int t = threadIdx.x;
  if (t == 0) bad = 0;
  __syncthreads();
  const int per = (ng + 1023) / 1024;
  const int lo = min(ng, t * per);
  const int hi = min(ng, lo + per);
  long long acc = 0;
  for (int i = lo; i < hi; ++i) {
    const int c = gptr[i];
    if (c < 0) bad = 1;
    acc += c;
  }
  sums[t] = acc;
  __syncthreads();
  // Inclusive scan of the 1024 segment sums (Hillis-Steele).
  for (int o = 1; o < 1024; o <<= 1) {
    const long long v = t >= o ? sums[t - o] : 0;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  long long run = sums[t] - acc;
  for (int i = lo; i < hi; ++i) {
    const int c = gptr[i];
    gptr[i] = (int)run;
    run += c;
  }
  if (t == 1023) {
    const long long total = sums[1023];
    gptr[ng] = (bad || total > INT_MAX) ? -1 : (int)total;
  }
}

// ---------------------------------------------------------------------------
// Product Y = A X for K right-hand sides (row-major [n][K]).
// A group team has K/CPL lanes per union-column slot (each lane owns CPL
// consecutive right-hand sides, one 16-byte X load) and NZ slots; a lane
// accumulates its CPL columns for all 4 rows of the group. Blocks take
// XCD-aware chunks of groups holding equal shares of the union columns.
// partial[block][c] receives the block's sum of X.Y (the p.q of CG).
// ---------------------------------------------------------------------------

template <typename MT>
__device__ inline void load_rb_vals(const MT* __restrict__ p, MT* v);
template <>
__device__ inline void load_rb_vals<float>(const float* __restrict__ p, float* v) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
}
template <>
__device__ inline void load_rb_vals<double>(const double* __restrict__ p, double* v) {
  const double2 a = *reinterpret_cast<const double2*>(p);
  const double2 b = *(reinterpret_cast<const double2*>(p) + 1);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = b.x;
  v[3] = b.y;
}

// Value r of the 4 held by the lanes of this lane's quad (DPP quad
// broadcast; no memory access).
template <int R>
__device__ inline float quad_bcast(float v) {
  constexpr int ctrl = R | (R << 2) | (R << 4) | (R << 6);
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), ctrl, 0xF, 0xF, false));
}

// RUNS: the union columns come from their runs (grptr, grun: first column,
// offset << 8 | length per run of consecutive columns, see runs.hip) in
// place of gcol. A team is then one wave (T = 64); per group it expands the
// runs into the wave's LDS slice (in chunks of kRbChunk columns) and the
// slots read their columns from there. The values and the summation order
// are those of the gcol product.
constexpr int kRbChunk = 1024;

template <typename MT, typename VT, int K, int NZ, int CPL, int UNR, bool RUNS>
__global__ __launch_bounds__(256) void k_spmm_rb(int n, int ng, const int* __restrict__ gptr,
                                                 const int* __restrict__ gcol, const MT* __restrict__ gval,
                                                 const VT* __restrict__ X, VT* __restrict__ Y,
                                                 double* __restrict__ partial, const ColState* __restrict__ st,
                                                 const int* __restrict__ grptr, const uint32_t* __restrict__ grun,
                                                 const VT* __restrict__ B) {
  constexpr int KL = K / CPL;
  constexpr int T = KL * NZ;
  constexpr int TPW = 64 / T;
  static_assert(!RUNS || T == 64, "the run-driven product takes one wave per group");
  // float32 values, 4 right-hand sides per lane and 4 lanes per slot.
  constexpr bool kQuadVals = sizeof(MT) == 4 && CPL == 4 && KL == 4;
  using AT = typename Accum<VT>::type;
  __shared__ double smem[4 * 16];
  __shared__ int s_rows[2];
  __shared__ int colbuf[RUNS ? 4 : 1][RUNS ? kRbChunk : 1];
  if (all_stopped<K>(st)) return;
  const int lane = threadIdx.x & 63;
  const int c0 = (lane % KL) * CPL;
  const int zs = (lane / KL) % NZ;
  const int wave = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int tpb = wpb * TPW;
  const int team = wave * TPW + lane / T;
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int chunk = (G % kXcds == 0) ? (b % kXcds) * (G / kXcds) + b / kXcds : b;
  if (wave == 0) {
    const long long total = gptr[ng];
    const int a0 = chunk == 0 ? 0 : row_lower_bound(gptr, ng, total * chunk / G);
    const int a1 = chunk == G - 1 ? ng : row_lower_bound(gptr, ng, total * (chunk + 1) / G);
    if (lane == 0) {
      s_rows[0] = a0;
      s_rows[1] = a1;
    }
  }
  __syncthreads();
  const int g0 = s_rows[0];
  const int g1 = s_rows[1];
  double dot[CPL];
#pragma unroll
  for (int m = 0; m < CPL; ++m) dot[m] = 0.0;
  for (int g = g0 + team; g < g1; g += tpb) {
    const int beg = gptr[g];
    const int len = gptr[g + 1] - beg;
    const MT* __restrict__ vp = gval + (size_t)beg * kRb;
    AT acc[kRb][CPL];
#pragma unroll
    for (int r = 0; r < kRb; ++r)
#pragma unroll
      for (int m = 0; m < CPL; ++m) acc[r][m] = 0;
    // Union columns [ja, jb) with column j at cp[j].
    auto columns = [&](const int* cp, int ja, int jb) {
      int j = ja + zs;
      if constexpr (kQuadVals && UNR == 4) {
        // The 4 lanes of a slot form a DPP quad: lane i of the quad loads
        // row value i of each column (the team's loads cover its columns'
        // values once, contiguously) and the quad shares them.
        const int qi = lane & 3;
        for (; j + 3 * NZ < jb; j += 4 * NZ) {
          int q[4];
          float vq[4];
          VT x[4][CPL];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            q[u] = cp[j + u * NZ];
            vq[u] = (float)vp[(size_t)(j + u * NZ) * kRb + qi];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) vload<VT, CPL>(X + (size_t)q[u] * K + c0, x[u]);
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const float vr[kRb] = {quad_bcast<0>(vq[u]), quad_bcast<1>(vq[u]), quad_bcast<2>(vq[u]),
                                   quad_bcast<3>(vq[u])};
#pragma unroll
            for (int r = 0; r < kRb; ++r)
#pragma unroll
              for (int m = 0; m < CPL; ++m) acc[r][m] += (AT)vr[r] * (AT)x[u][m];
          }
        }
      } else if constexpr (UNR == 4) {
        for (; j + 3 * NZ < jb; j += 4 * NZ) {
          int q[4];
          MT v[4][kRb];
          VT x[4][CPL];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            q[u] = cp[j + u * NZ];
            load_rb_vals<MT>(vp + (size_t)(j + u * NZ) * kRb, v[u]);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) vload<VT, CPL>(X + (size_t)q[u] * K + c0, x[u]);
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int r = 0; r < kRb; ++r)
#pragma unroll
              for (int m = 0; m < CPL; ++m) acc[r][m] += (AT)v[u][r] * (AT)x[u][m];
        }
      }
      for (; j < jb; j += NZ) {
        MT v[kRb];
        VT x[CPL];
        load_rb_vals<MT>(vp + (size_t)j * kRb, v);
        vload<VT, CPL>(X + (size_t)cp[j] * K + c0, x);
#pragma unroll
        for (int r = 0; r < kRb; ++r)
#pragma unroll
          for (int m = 0; m < CPL; ++m) acc[r][m] += (AT)v[r] * (AT)x[m];
      }
    };
    if constexpr (RUNS) {
      int* slice = colbuf[wave];
      const int rb = grptr[g], re = grptr[g + 1];
      for (int a = 0; a < len; a += kRbChunk) {
        const int b = min(len, a + kRbChunk);
        for (int k = rb + lane; k < re; k += 64) {
          const uint2 rr = *reinterpret_cast<const uint2*>(grun + 2 * (size_t)k);
          const int off = (int)(rr.y >> 8), ln = (int)(rr.y & 255u);
          const int lo = max(off, a), hi = min(off + ln, b);
          for (int e = lo; e < hi; ++e) slice[e - a] = (int)rr.x + (e - off);
        }
        // The slice is written and read by this wave only; its LDS accesses
        // complete in order.
        __builtin_amdgcn_wave_barrier();
        columns(slice - a, a, b);
        __builtin_amdgcn_wave_barrier();
      }
    } else {
      columns(gcol + beg, 0, len);
    }
#pragma unroll
    for (int o = KL; o < T; o <<= 1)
#pragma unroll
      for (int r = 0; r < kRb; ++r)
#pragma unroll
        for (int m = 0; m < CPL; ++m) acc[r][m] += __shfl_xor(acc[r][m], o, 64);
    if (zs == 0) {
#pragma unroll
      for (int r = 0; r < kRb; ++r) {
        const int row = g * kRb + r;
        if (row < n) {
          const size_t e = (size_t)row * K + c0;
          VT yv[CPL];
          if (B) {
            // Residual form: Y = B - A X, partials Y.Y.
            VT bv[CPL];
            vload<VT, CPL>(B + e, bv);
#pragma unroll
            for (int m = 0; m < CPL; ++m) yv[m] = (VT)((AT)bv[m] - acc[r][m]);
            vstore<VT, CPL>(Y + e, yv);
#pragma unroll
            for (int m = 0; m < CPL; ++m) dot[m] += (double)yv[m] * (double)yv[m];
          } else {
#pragma unroll
            for (int m = 0; m < CPL; ++m) yv[m] = (VT)acc[r][m];
            vstore<VT, CPL>(Y + e, yv);
            if (partial) {
              VT xv[CPL];
              vload<VT, CPL>(X + e, xv);
#pragma unroll
              for (int m = 0; m < CPL; ++m) dot[m] += (double)xv[m] * (double)yv[m];
            }
          }
        }
      }
    }
  }
  if (partial) {
#pragma unroll
    for (int o = KL; o < 64; o <<= 1)
#pragma unroll
      for (int m = 0; m < CPL; ++m) dot[m] += __shfl_xor(dot[m], o, 64);
    if (lane < KL)
#pragma unroll
      for (int m = 0; m < CPL; ++m) smem[wave * K + c0 + m] = dot[m];
    __syncthreads();
    if (threadIdx.x < K) {
      double s = 0.0;
      for (int w = 0; w < wpb; ++w) s += smem[w * K + threadIdx.x];
      partial[(size_t)b * K + threadIdx.x] = s;
    }
  }
}

// The product from paired row blocks: as k_spmm_rb with 8 rows per group
// (a lane accumulates its CPL columns for all 8 rows) and the rows of group
// g taken from the pair list.
template <typename MT, typename VT, int K, int NZ, int CPL, int UNR>
__global__ __launch_bounds__(256) void k_spmm_rb8(int n, int ng, const int* __restrict__ pairs,
                                                  const int* __restrict__ gptr, const int* __restrict__ gcol,
                                                  const MT* __restrict__ gval, const VT* __restrict__ X,
                                                  VT* __restrict__ Y, double* __restrict__ partial,
                                                  const ColState* __restrict__ st) {
  constexpr int KL = K / CPL;
  constexpr int T = KL * NZ;
  constexpr int TPW = 64 / T;
  using AT = typename Accum<VT>::type;
  __shared__ double smem[4 * 16];
  __shared__ int s_rows[2];
  if (all_stopped<K>(st)) return;
  const int lane = threadIdx.x & 63;
  const int c0 = (lane % KL) * CPL;
  const int zs = (lane / KL) % NZ;
  const int wave = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int tpb = wpb * TPW;
  const int team = wave * TPW + lane / T;
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int chunk = (G % kXcds == 0) ? (b % kXcds) * (G / kXcds) + b / kXcds : b;
  if (wave == 0) {
    const long long total = gptr[ng];
    const int a0 = chunk == 0 ? 0 : row_lower_bound(gptr, ng, total * chunk / G);
    const int a1 = chunk == G - 1 ? ng : row_lower_bound(gptr, ng, total * (chunk + 1) / G);
    if (lane == 0) {
      s_rows[0] = a0;
      s_rows[1] = a1;
    }
  }
  __syncthreads();
  const int g0 = s_rows[0];
  const int g1 = s_rows[1];
  double dot[CPL];
#pragma unroll
  for (int m = 0; m < CPL; ++m) dot[m] = 0.0;
  for (int g = g0 + team; g < g1; g += tpb) {
    const int beg = gptr[g];
    const int len = gptr[g + 1] - beg;
    const int* __restrict__ cp = gcol + beg;
    const MT* __restrict__ vp = gval + (size_t)beg * kRb2;
    AT acc[kRb2][CPL];
#pragma unroll
    for (int r = 0; r < kRb2; ++r)
#pragma unroll
      for (int m = 0; m < CPL; ++m) acc[r][m] = 0;
    int j = zs;
    if constexpr (UNR == 4) {
      for (; j + 3 * NZ < len; j += 4 * NZ) {
        int q[4];
        MT v[4][kRb2];
        VT x[4][CPL];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          q[u] = cp[j + u * NZ];
          load_rb_vals<MT>(vp + (size_t)(j + u * NZ) * kRb2, v[u]);
          load_rb_vals<MT>(vp + (size_t)(j + u * NZ) * kRb2 + kRb, v[u] + kRb);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) vload<VT, CPL>(X + (size_t)q[u] * K + c0, x[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int r = 0; r < kRb2; ++r)
#pragma unroll
            for (int m = 0; m < CPL; ++m) acc[r][m] += (AT)v[u][r] * (AT)x[u][m];
      }
    }
    for (; j < len; j += NZ) {
      MT v[kRb2];
      VT x[CPL];
      load_rb_vals<MT>(vp + (size_t)j * kRb2, v);
      load_rb_vals<MT>(vp + (size_t)j * kRb2 + kRb, v + kRb);
      vload<VT, CPL>(X + (size_t)cp[j] * K + c0, x);
#pragma unroll
      for (int r = 0; r < kRb2; ++r)
#pragma unroll
        for (int m = 0; m < CPL; ++m) acc[r][m] += (AT)v[r] * (AT)x[m];
    }
#pragma unroll
    for (int o = KL; o < T; o <<= 1)
#pragma unroll
      for (int r = 0; r < kRb2; ++r)
#pragma unroll
        for (int m = 0; m < CPL; ++m) acc[r][m] += __shfl_xor(acc[r][m], o, 64);
    if (zs == 0) {
#pragma unroll
      for (int r = 0; r < kRb2; ++r) {
        const int row = pair_row(pairs, g, r, n);
        if (row >= 0) {
          const size_t e = (size_t)row * K + c0;
          VT yv[CPL];
#pragma unroll
          for (int m = 0; m < CPL; ++m) yv[m] = (VT)acc[r][m];
          vstore<VT, CPL>(Y + e, yv);
          if (partial) {
            VT xv[CPL];
            vload<VT, CPL>(X + e, xv);
#pragma unroll
            for (int m = 0; m < CPL; ++m) dot[m] += (double)xv[m] * (double)yv[m];
          }
        }
      }
    }
  }
  if (partial) {
#pragma unroll
    for (int o = KL; o < 64; o <<= 1)
#pragma unroll
      for (int m = 0; m < CPL; ++m) dot[m] += __shfl_xor(dot[m], o, 64);
    if (lane < KL)
#pragma unroll
      for (int m = 0; m < CPL; ++m) smem[wave * K + c0 + m] = dot[m];
    __syncthreads();
    if (threadIdx.x < K) {
      double s = 0.0;
      for (int w = 0; w < wpb; ++w) s += smem[w * K + threadIdx.x];
      partial[(size_t)b * K + threadIdx.x] = s;
    }
  }
}

// Launch shape per value width and k (set by the autotuner; 0 = default).
struct RbConfig {
  int nz;      // union-column slots per group team ((k/cpl)*nz <= 64)
  int grid;    // blocks (<= kMaxParts)
  int unroll;  // 1 or 4 columns in flight per lane
  int cpl;     // right-hand sides per lane (16-byte loads at most)
};
static RbConfig g_rb_cfg[2][17] = {};

template <typename MT, typename VT, int K, int NZ, int CPL>
static int rb_launch_nz(const cc_rb_desc& rb, const MT* gval, const VT* X, VT* Y, double* partial, const ColState* st,
                        int grid_req, int unroll, hipStream_t s, int* grid_out, const VT* B) {
  constexpr int T = (K / CPL) * NZ;
  const int teams_per_block = (kBlock / 64) * (64 / T);
  int grid = grid_req;
  if (grid <= 0) grid = (rb.ngroups + teams_per_block - 1) / teams_per_block;
  if (grid > kMaxParts) grid = kMaxParts;
  if (grid >= 2 * kXcds) grid -= grid % kXcds;
  if (grid < 1) grid = 1;
  constexpr bool kTeamIsWave = T == 64;
  if (rb.pairs) {
    if (B) return CC_E_ARG;  // the paired form has no residual form
    if (unroll == 4)
      launch_timed(k_spmm_rb8<MT, VT, K, NZ, CPL, 4>, dim3(grid), dim3(kBlock), 0, s, rb.n, rb.ngroups,
                   rb.pairs, rb.gptr, rb.gcol, gval, X, Y, partial, st);
    else
      launch_timed(k_spmm_rb8<MT, VT, K, NZ, CPL, 1>, dim3(grid), dim3(kBlock), 0, s, rb.n, rb.ngroups,
                   rb.pairs, rb.gptr, rb.gcol, gval, X, Y, partial, st);
  } else if (kTeamIsWave && rb.grun && rb.grptr) {
    if (unroll == 4)
      launch_timed(k_spmm_rb<MT, VT, K, NZ, CPL, 4, kTeamIsWave>, dim3(grid), dim3(kBlock), 0, s, rb.n,
                   rb.ngroups, rb.gptr, rb.gcol, gval, X, Y, partial, st, rb.grptr, rb.grun, B);
    else
      launch_timed(k_spmm_rb<MT, VT, K, NZ, CPL, 1, kTeamIsWave>, dim3(grid), dim3(kBlock), 0, s, rb.n,
                   rb.ngroups, rb.gptr, rb.gcol, gval, X, Y, partial, st, rb.grptr, rb.grun, B);
  } else if (unroll == 4) {
    launch_timed(k_spmm_rb<MT, VT, K, NZ, CPL, 4, false>, dim3(grid), dim3(kBlock), 0, s, rb.n, rb.ngroups,
                 rb.gptr, rb.gcol, gval, X, Y, partial, st, nullptr, nullptr, B);
  } else {
    launch_timed(k_spmm_rb<MT, VT, K, NZ, CPL, 1, false>, dim3(grid), dim3(kBlock), 0, s, rb.n, rb.ngroups,
                 rb.gptr, rb.gcol, gval, X, Y, partial, st, nullptr, nullptr, B);
  }
  if (grid_out) *grid_out = grid;
  return (int)hipGetLastError();
}

template <typename MT, typename VT, int K, int CPL>
static int rb_launch_cpl(const cc_rb_desc& rb, const MT* gval, const VT* X, VT* Y, double* partial,
                         const ColState* st, const RbConfig& cfg, hipStream_t s, int* grid_out, const VT* B) {
  constexpr int KL = K / CPL;
  // Default: 64-lane teams, at most 16 slots.
  int nz = cfg.nz > 0 && KL * cfg.nz <= 64 ? cfg.nz : (64 / KL < 16 ? 64 / KL : 16);
  const int unroll = cfg.unroll == 1 ? 1 : 4;
#define CC_RB_NZ(NZV)                                                                                         \
  if (nz == NZV && KL * NZV <= 64)                                                                            \
    return rb_launch_nz<MT, VT, K, (KL * NZV <= 64 ? NZV : 1), CPL>(rb, gval, X, Y, partial, st, cfg.grid, \
                                                                    unroll, s, grid_out, B);
  CC_RB_NZ(64)
  CC_RB_NZ(32)
  CC_RB_NZ(16)
  CC_RB_NZ(8)
  CC_RB_NZ(4)
  CC_RB_NZ(2)
  CC_RB_NZ(1)
#undef CC_RB_NZ
  return CC_E_ARG;
}

template <typename MT, typename VT, int K>
int rb_spmm(const cc_rb_desc& rb, const VT* X, VT* Y, double* partial, const ColState* st, hipStream_t s,
            int* grid_out, const VT* B) {
  const MT* gval = nullptr;
  if constexpr (sizeof(MT) == 8)
    gval = rb.gval;
  else
    gval = rb.gval32;
  if (!gval) return CC_E_ARG;
  const RbConfig& cfg = g_rb_cfg[sizeof(VT) == 4 ? 1 : 0][K];
  constexpr int wide = 16 / (int)sizeof(VT);
  int cpl = cfg.cpl > 0 ? cfg.cpl : (K >= wide ? wide : K);
  if (cpl > K || cpl * (int)sizeof(VT) > 16) cpl = K >= wide ? wide : K;
  if (cpl >= 4 && sizeof(VT) == 4 && K >= 4)
    return rb_launch_cpl<MT, VT, K, (sizeof(VT) == 4 && K >= 4 ? 4 : 1)>(rb, gval, X, Y, partial, st, cfg, s,
                                                                        grid_out, B);
  if (cpl >= 2 && K >= 2)
    return rb_launch_cpl<MT, VT, K, (K >= 2 ? 2 : 1)>(rb, gval, X, Y, partial, st, cfg, s, grid_out, B);
  return rb_launch_cpl<MT, VT, K, 1>(rb, gval, X, Y, partial, st, cfg, s, grid_out, B);
}

#define CC_RB_INST(K)                                                                                          \
  template int rb_spmm<double, double, K>(const cc_rb_desc&, const double*, double*, double*, const ColState*, \
                                          hipStream_t, int*, const double*);                                   \
  template int rb_spmm<float, float, K>(const cc_rb_desc&, const float*, float*, double*, const ColState*,     \
                                        hipStream_t, int*, const float*);
CC_RB_INST(1)
CC_RB_INST(2)
CC_RB_INST(4)
CC_RB_INST(8)
CC_RB_INST(16)
#undef CC_RB_INST

bool rb_desc_ok(const cc_rb_desc* rb, int n) {
  if (!rb || rb->n != n || !rb->gptr || !rb->gcol) return false;
  if (rb->pairs) return rb->ngroups >= 1 && rb->ngroups <= (n + kRb - 1) / kRb && !rb->grun;
  return rb->ngroups == (n + kRb - 1) / kRb;
}

}  // namespace cc

using namespace cc;

extern "C" {

int cc_rb_build_count(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* gptr, int64_t* nu,
                      cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !gptr || !nu) return CC_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int ng = (n + kRb - 1) / kRb;
  hipLaunchKernelGGL((k_rb_merge<false>), dim3((ng + 255) / 256), dim3(256), 0, s, n, ng, indptr, indices, nullptr,
                     gptr, nullptr, nullptr, nullptr);
  hipLaunchKernelGGL(k_rb_scan, dim3(1), dim3(1024), 0, s, ng, gptr);
  CC_CHECK(hipGetLastError());
  int32_t total = 0;
  CC_CHECK(hipMemcpyAsync(&total, gptr + ng, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  if (total < 0) return CC_E_ARG;
  *nu = total;
  return 0;
}

int cc_rb_build_fill(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                     const int32_t* gptr, int32_t* gcol, double* gval, float* gval32, cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !vals || !gptr || !gcol || (!gval && !gval32)) return CC_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int ng = (n + kRb - 1) / kRb;
  hipLaunchKernelGGL((k_rb_merge<true>), dim3((ng + 255) / 256), dim3(256), 0, s, n, ng, indptr, indices, vals,
                     const_cast<int32_t*>(gptr), gcol, gval, gval32);
  return (int)hipGetLastError();
}

int cc_rb_build_pairs_count(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t npairs,
                            const int32_t* pairs, int32_t* gptr, int64_t* nu, cc_stream_t stream) {
  if (n <= 0 || npairs <= 0 || npairs > (n + kRb - 1) / kRb || !indptr || !indices || !pairs || !gptr || !nu)
    return CC_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL((k_rb8_merge<false>), dim3((npairs + 255) / 256), dim3(256), 0, s, n, npairs, pairs, indptr,
                     indices, nullptr, gptr, nullptr, nullptr, nullptr);
  hipLaunchKernelGGL(k_rb_scan, dim3(1), dim3(1024), 0, s, npairs, gptr);
  CC_CHECK(hipGetLastError());
  int32_t total = 0;
  CC_CHECK(hipMemcpyAsync(&total, gptr + npairs, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  if (total < 0) return CC_E_ARG;
  *nu = total;
  return 0;
}

int cc_rb_build_pairs_fill(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                           int32_t npairs, const int32_t* pairs, const int32_t* gptr, int32_t* gcol, double* gval,
                           float* gval32, cc_stream_t stream) {
  if (n <= 0 || npairs <= 0 || !indptr || !indices || !vals || !pairs || !gptr || !gcol || (!gval && !gval32))
    return CC_E_ARG;
  hipLaunchKernelGGL((k_rb8_merge<true>), dim3((npairs + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, npairs,
                     pairs, indptr, indices, vals, const_cast<int32_t*>(gptr), gcol, gval, gval32);
  return (int)hipGetLastError();
}

int cc_rb_set_config(int32_t k, int32_t value_bytes, int32_t nz_lanes, int32_t grid, int32_t unroll,
                     int32_t cols_per_lane) {
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16) return CC_E_ARG;
  if (value_bytes != 4 && value_bytes != 8) return CC_E_ARG;
  if (cols_per_lane != 0 && cols_per_lane != 1 && cols_per_lane != 2 && cols_per_lane != 4) return CC_E_ARG;
  if (cols_per_lane * value_bytes > 16 || cols_per_lane > k) return CC_E_ARG;
  const int kl = k / (cols_per_lane > 0 ? cols_per_lane : 1);
  if (nz_lanes < 0 || kl * nz_lanes > 64 || (nz_lanes & (nz_lanes - 1)) != 0) return CC_E_ARG;
  if (grid < 0 || grid > kMaxParts || (unroll != 0 && unroll != 1 && unroll != 4)) return CC_E_ARG;
  g_rb_cfg[value_bytes == 4 ? 1 : 0][k] = RbConfig{nz_lanes, grid, unroll, cols_per_lane};
  return 0;
}

int cc_spmm_rb_f64(const cc_rb_desc* rb, int32_t k, const double* X, double* Y, cc_stream_t stream) {
  if (!rb || !rb_desc_ok(rb, rb->n) || !rb->gval || !X || !Y) return CC_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  switch (k) {
    case 1: return rb_spmm<double, double, 1>(*rb, X, Y, nullptr, nullptr, s, nullptr);
    case 2: return rb_spmm<double, double, 2>(*rb, X, Y, nullptr, nullptr, s, nullptr);
    case 4: return rb_spmm<double, double, 4>(*rb, X, Y, nullptr, nullptr, s, nullptr);
    case 8: return rb_spmm<double, double, 8>(*rb, X, Y, nullptr, nullptr, s, nullptr);
    case 16: return rb_spmm<double, double, 16>(*rb, X, Y, nullptr, nullptr, s, nullptr);
    default: return CC_E_ARG;
  }
}

int cc_spmm_rb_f32(const cc_rb_desc* rb, int32_t k, const float* X, float* Y, cc_stream_t stream) {
  if (!rb || !rb_desc_ok(rb, rb->n) || !rb->gval32 || !X || !Y) return CC_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  switch (k) {
    case 1: return rb_spmm<float, float, 1>(*rb, X, Y, nullptr, nullptr, s, nullptr);
    case 2: return rb_spmm<float, float, 2>(*rb, X, Y, nullptr, nullptr, s, nullptr);
    case 4: return rb_spmm<float, float, 4>(*rb, X, Y, nullptr, nullptr, s, nullptr);
    case 8: return rb_spmm<float, float, 8>(*rb, X, Y, nullptr, nullptr, s, nullptr);
    case 16: return rb_spmm<float, float, 16>(*rb, X, Y, nullptr, nullptr, s, nullptr);
    default: return CC_E_ARG;
  }
}

}  // extern "C"
