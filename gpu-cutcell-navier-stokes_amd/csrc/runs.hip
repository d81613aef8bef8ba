// runs.hip - run encoding of sorted column lists and the single-vector
// product over it.
//
// On the 3D B-spline systems a row's columns come in runs of consecutive
// indices (the stencil's x-lines: ~7 per run on the p = 3 planar system,
// 49 runs for an interior row of 343 nonzeros). A unit (a CSR row, or a
// row-block group's union) is stored as its runs, 8 bytes each:
//     runs[2 r]     = first column of run r
//     runs[2 r + 1] = (offset of the run's first entry within the unit) << 8
//                     | run length (1 .. 255)
// in place of one 4-byte index per entry: on the 81k-unknown planar system
// 3.95M runs replace 27.4M indices, so a float64 SpMV streams 9.2 instead of
// 12 bytes per nonzero and a float32 one 5.2 instead of 8.
//
// Build: a count pass (one thread per unit), a one-block scan to run
// offsets, a fill pass; all on the device.

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "../../include/cutcell_hip.h"
#include "common.h"
#include "spmm_util.h"

namespace cc {

constexpr int kRunMax = 255;

// Runs of unit u (entries uptr[u] .. uptr[u+1]-1 of cols). FILL = false:
// count[u] = number of runs (-1 if the columns do not ascend strictly or the
// unit is too long for 24-bit offsets). FILL = true: write the runs at
// rptr[u].
template <bool FILL>
__global__ __launch_bounds__(256) void k_runs_scan_units(int nunits, const int* __restrict__ uptr,
                                                         const int* __restrict__ cols, int* __restrict__ rptr,
                                                         uint32_t* __restrict__ runs) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nunits) return;
  const int b = uptr[u], e = uptr[u + 1];
  int out = FILL ? rptr[u] : 0;
  bool bad = (e - b) >= (1 << 24);
  int start = 0, len = 0, off = 0, prev = 0;
  for (int i = b; i < e; ++i) {
    const int c = cols[i];
    if (len > 0 && c <= prev) bad = true;
    if (len > 0 && c == prev + 1 && len < kRunMax) {
      ++len;
    } else {
      if (len > 0) {
        if (FILL) {
          runs[2 * (size_t)out] = (uint32_t)start;
          runs[2 * (size_t)out + 1] = ((uint32_t)off << 8) | (uint32_t)len;
        }
        ++out;
      }
      start = c;
      len = 1;
      off = i - b;
    }
    prev = c;
  }
  if (len > 0) {
    if (FILL) {
      runs[2 * (size_t)out] = (uint32_t)start;
      runs[2 * (size_t)out + 1] = ((uint32_t)off << 8) | (uint32_t)len;
    }
    ++out;
  }
  if (!FILL) rptr[u] = bad ? -1 : out;
}

// One block of 1024 threads: counts rptr[0, nunits) -> exclusive offsets,
// rptr[nunits] = total, or -1 if a count was negative or the total does not
// fit an int32.
__global__ __launch_bounds__(1024) void k_runs_offsets(int nunits, int* __restrict__ rptr) {
  __shared__ long long sums[1024];
  __shared__ int bad;
  const int t = threadIdx.x;
  if (t == 0) bad = 0;
  __syncthreads();
  const int per = (nunits + 1023) / 1024;
  const int lo = min(nunits, t * per);
  const int hi = min(nunits, lo + per);
  long long acc = 0;
  for (int i = lo; i < hi; ++i) {
    const int c = rptr[i];
    if (c < 0) bad = 1;
    acc += c;
  }
  sums[t] = acc;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const long long v = t >= o ? sums[t - o] : 0;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  long long run = sums[t] - acc;
  for (int i = lo; i < hi; ++i) {
    const int c = rptr[i];
    rptr[i] = (int)run;
    run += c;
  }
  if (t == 1023) {
    const long long total = sums[1023];
    rptr[nunits] = (bad || total > INT_MAX) ? -1 : (int)total;
  }
}

// y = A x from the runs of the rows: one wave per row. The wave expands the
// row's runs into a column list in its own LDS slice (one run per lane,
// entries written in order), then streams the row's values as CSR does, lane
// by entry (coalesced loads), with each entry's column read from LDS.
// A row of at most kRunFast entries and 64 runs (every interior row of the
// B-spline systems) takes one pass: its values are loaded before the runs
// are expanded (they do not depend on them) and the next row's pointers are
// loaded during this one. Longer rows go in chunks of the slice. Each lane
// sums its entries in order; the wave then sums its lanes in a fixed
// butterfly order.
constexpr int kRunChunk = 1024;  // columns per wave slice (4 KB of LDS)
constexpr int kRunWaves = 4;     // waves per block
constexpr int kRunFast = 512;    // entries of a one-pass row (8 per lane)

// partial (may be NULL): per-block sum of X[r] Y[r] over the block's rows
// (the p.q of CG), waves added in order; st (may be NULL): skip the launch
// once the column has stopped.
template <typename MT, typename VT>
__global__ __launch_bounds__(64 * kRunWaves) void k_spmv_runs(int n, const int* __restrict__ rptr,
                                                              const uint32_t* __restrict__ runs,
                                                              const int* __restrict__ indptr,
                                                              const MT* __restrict__ vals,
                                                              const VT* __restrict__ X, VT* __restrict__ Y,
                                                              double* __restrict__ partial,
                                                              const ColState* __restrict__ st) {
  using AT = typename Accum<VT>::type;
  __shared__ int colbuf[kRunWaves][kRunChunk];
  __shared__ double s_dot[kRunWaves];
  if (all_stopped<1>(st)) return;
  double dot = 0.0;
  const int lane = threadIdx.x & 63;
  const int wl = threadIdx.x >> 6;
  int* cols = colbuf[wl];
  const int w0 = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nw = (int)((gridDim.x * blockDim.x) >> 6);
  int r = w0;
  int rb = 0, re = 0, vb = 0, ve = 0;
  if (r < n) {
    rb = rptr[r];
    re = rptr[r + 1];
    vb = indptr[r];
    ve = indptr[r + 1];
  }
  for (; r < n; r += nw) {
    const int len = ve - vb;
    const int nrun = re - rb;
    const int cr = rb, cv = vb;
    // The next row's pointers.
    const int rn = r + nw;
    if (rn < n) {
      rb = rptr[rn];
      re = rptr[rn + 1];
      vb = indptr[rn];
      ve = indptr[rn + 1];
    }
    AT acc = 0;
    const MT* __restrict__ vr = vals + cv;
    if (len <= kRunFast && nrun <= 64) {
      MT v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = lane + 64 * u;
        v[u] = e < len ? vr[e] : (MT)0;
      }
      if (lane < nrun) {
        const uint2 rr = *reinterpret_cast<const uint2*>(runs + 2 * (size_t)(cr + lane));
        const int off = (int)(rr.y >> 8), ln = (int)(rr.y & 255u);
        for (int i = 0; i < ln; ++i) cols[off + i] = (int)rr.x + i;
      }
      // The slice is written and read by this wave only; its LDS accesses
      // complete in order.
      __builtin_amdgcn_wave_barrier();
      VT x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = lane + 64 * u;
        x[u] = e < len ? X[cols[e]] : (VT)0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += (AT)v[u] * (AT)x[u];
      __builtin_amdgcn_wave_barrier();
    } else {
      for (int c0 = 0; c0 < len; c0 += kRunChunk) {
        const int c1 = min(len, c0 + kRunChunk);
        for (int k = cr + lane; k < cr + nrun; k += 64) {
          const uint2 rr = *reinterpret_cast<const uint2*>(runs + 2 * (size_t)k);
          const int off = (int)(rr.y >> 8), ln = (int)(rr.y & 255u);
          const int lo = max(off, c0), hi = min(off + ln, c1);
          for (int e = lo; e < hi; ++e) cols[e - c0] = (int)rr.x + (e - off);
        }
        __builtin_amdgcn_wave_barrier();
        for (int e = c0 + lane; e < c1; e += 64) acc += (AT)vr[e] * (AT)X[cols[e - c0]];
        __builtin_amdgcn_wave_barrier();
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) {
      const VT y = (VT)acc;
      Y[r] = y;
      if (partial) dot += (double)X[r] * (double)y;
    }
  }
  if (partial) {
    if (lane == 0) s_dot[wl] = dot;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kRunWaves; ++w) t += s_dot[w];
      partial[blockIdx.x] = t;
    }
  }
}

template <typename MT, typename VT>
static int spmv_runs_launch(int n, const int* rptr, const uint32_t* runs, const int* indptr, const MT* vals,
                            const VT* X, VT* Y, hipStream_t s, double* partial = nullptr,
                            const ColState* st = nullptr, int* grid_out = nullptr) {
  if (n == 0) return 0;
  // One row per wave on the first pass, at most 16 blocks per CU; with
  // partials at most kMaxParts blocks (one partial each).
  int grid = (n + kRunWaves - 1) / kRunWaves;
  const int cap = partial ? kMaxParts : 256 * 16;
  if (grid > cap) grid = cap;
  launch_timed(k_spmv_runs<MT, VT>, dim3(grid), dim3(64 * kRunWaves), 0, s, n, rptr, runs, indptr, vals, X,
               Y, partial, st);
  if (grid_out) *grid_out = grid;
  return (int)hipGetLastError();
}

// The CG product from the runs (pcg.hip): Q = A P with per-block P.Q.
template <typename MT, typename VT>
int runs_spmv_cg(int n, const int* rptr, const uint32_t* runs, const int* indptr, const MT* vals, const VT* P, VT* Q,
                 double* partial, const ColState* st, hipStream_t s, int* grid_out) {
  return spmv_runs_launch<MT, VT>(n, rptr, runs, indptr, vals, P, Q, s, partial, st, grid_out);
}
template int runs_spmv_cg<double, double>(int, const int*, const uint32_t*, const int*, const double*, const double*,
                                          double*, double*, const ColState*, hipStream_t, int*);
template int runs_spmv_cg<float, float>(int, const int*, const uint32_t*, const int*, const float*, const float*,
                                        float*, double*, const ColState*, hipStream_t, int*);

}  // namespace cc

using namespace cc;

extern "C" {

int cc_runs_count(int32_t nunits, const int32_t* uptr, const int32_t* cols, int32_t* rptr, int64_t* nruns,
                  cc_stream_t stream) {
  if (nunits < 0 || !nruns || (nunits > 0 && (!uptr || !cols || !rptr))) return CC_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  *nruns = 0;
  if (nunits == 0) return 0;
  hipLaunchKernelGGL(k_runs_scan_units<false>, dim3((nunits + 255) / 256), dim3(256), 0, s, nunits, uptr, cols, rptr,
                     nullptr);
  hipLaunchKernelGGL(k_runs_offsets, dim3(1), dim3(1024), 0, s, nunits, rptr);
  CC_CHECK(hipGetLastError());
  int32_t total = 0;
  CC_CHECK(hipMemcpyAsync(&total, rptr + nunits, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  if (total < 0) return CC_E_ARG;
  *nruns = total;
  return 0;
}

int cc_runs_fill(int32_t nunits, const int32_t* uptr, const int32_t* cols, const int32_t* rptr, uint32_t* runs,
                 cc_stream_t stream) {
  if (nunits < 0 || (nunits > 0 && (!uptr || !cols || !rptr || !runs))) return CC_E_ARG;
  if (nunits == 0) return 0;
  hipLaunchKernelGGL(k_runs_scan_units<true>, dim3((nunits + 255) / 256), dim3(256), 0, (hipStream_t)stream, nunits,
                     uptr, cols, const_cast<int32_t*>(rptr), runs);
  return (int)hipGetLastError();
}

int cc_spmv_runs_f64(int32_t n, const int32_t* rptr, const uint32_t* runs, const int32_t* indptr, const double* vals,
                     const double* x, double* y, cc_stream_t stream) {
  if (n < 0 || (n > 0 && (!rptr || !runs || !indptr || !vals || !x || !y))) return CC_E_ARG;
  return spmv_runs_launch<double, double>(n, rptr, runs, indptr, vals, x, y, (hipStream_t)stream);
}

int cc_spmv_runs_f32(int32_t n, const int32_t* rptr, const uint32_t* runs, const int32_t* indptr, const float* vals,
                     const float* x, float* y, cc_stream_t stream) {
  if (n < 0 || (n > 0 && (!rptr || !runs || !indptr || !vals || !x || !y))) return CC_E_ARG;
  return spmv_runs_launch<float, float>(n, rptr, runs, indptr, vals, x, y, (hipStream_t)stream);
}

}  // extern "C"
