// batched.hip - many independent cut-cell systems of uneven size, one launch.
//
// A persistent grid of workgroups pulls systems from a device work counter
// (systems pre-sorted longest-first by the host). One workgroup runs the whole
// Jacobi-PCG loop of its system: the search direction p and the product q live
// in LDS (the system's vectors fit: n <= 8192 rows in float64), x, r, z and
// the diagonal live in registers of the threads that own the rows, and every
// dot product is a deterministic wave-shuffle + LDS reduction inside the
// workgroup, so no iteration ever leaves the CU. The loop semantics are those
// of the reference GPU loop (code/fast_solver_backends.py:233-274), applied
// per system; the float64 true residual is recomputed on the device at exit.
//
// MIXED = true: float32 values and vectors in an inner Jacobi-PCG, float64
// right-hand side, solution, residual recomputation (with the float64 matrix)
// and correction in an outer refinement loop, as in the single-system mixed
// solver.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {

template <int NW>
__device__ inline double wg_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  __syncthreads();  // red may still be read by the previous reduction
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) s += red[w];
  return s;
}

template <int NW>
__device__ inline void wg_sum2(double& a, double& b, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[wave] = a;
    red[NW + wave] = b;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    sa += red[w];
    sb += red[NW + w];
  }
  a = sa;
  b = sb;
}

// Team SpMV inside one workgroup: y[row] = sum_j vals[j] * xin[indices[j]]
// with xin in LDS. Returns this thread's share of sum_row w[row] * y[row]
// (w = xin for p.Ap, or the residual mode's r[row]^2 by the caller).
// MODE 0: y = A x (written to LDS), dot += x[row] * y[row]
// MODE 1: y = b - A x (written to LDS), dot += y[row]^2
template <typename MT, typename VT, int T, int MODE>
__device__ inline double team_spmv(int n, const int* __restrict__ ip, const int* __restrict__ ix,
                                   const MT* __restrict__ va, const VT* xin, VT* yout,
                                   const double* __restrict__ b) {
  const int lane = threadIdx.x & 63;
  const int tl = lane % T;
  const int team = threadIdx.x / T;
  const int nteams = blockDim.x / T;
  double dot = 0.0;
  // R rows per team at a time (row, row + nteams, ...), their entries eight
  // per lane: all indices and values are loaded before the gathers, instead
  // of one dependent round trip per entry. Each row's sum keeps the entry
  // order and the rows are finished in order, so the results are those of
  // the one-entry-at-a-time loop. Two rows for float32 values; one for
  // float64, whose loads in flight would cost occupancy (measured slower).
  constexpr int R = sizeof(MT) == 4 ? 2 : 1;
  for (int row = team; row < n; row += R * nteams) {
    int rows[R];
    VT acc[R];
    int beg[R], end[R];
#pragma unroll
    for (int h = 0; h < R; ++h) {
      rows[h] = row + h * nteams;
      acc[h] = (VT)0;
    }
#pragma unroll
    for (int h = 0; h < R; ++h) {
      const bool live = rows[h] < n;
      beg[h] = live ? ip[rows[h]] : 0;
      end[h] = live ? ip[rows[h] + 1] : 0;
    }
    int span = end[0] - beg[0];
#pragma unroll
    for (int h = 1; h < R; ++h) span = max(span, end[h] - beg[h]);
    for (int o0 = tl; o0 < span; o0 += 8 * T) {
      int c[R][8];
      MT v[R][8];
#pragma unroll
      for (int h = 0; h < R; ++h)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int j = beg[h] + o0 + u * T;
          const bool ok = j < end[h];
          c[h][u] = ok ? ix[j] : 0;
          v[h][u] = ok ? va[j] : (MT)0;
        }
#pragma unroll
      for (int h = 0; h < R; ++h)
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (beg[h] + o0 + u * T < end[h]) acc[h] += (VT)v[h][u] * xin[c[h][u]];
    }
#pragma unroll
    for (int h = 0; h < R; ++h) {
#pragma unroll
      for (int o = 1; o < T; o <<= 1) acc[h] += __shfl_xor(acc[h], o, 64);
    }
#pragma unroll
    for (int h = 0; h < R; ++h) {
      const int r = rows[h];
      if (tl == 0 && r < n) {
        if (MODE == 0) {
          yout[r] = acc[h];
          dot += (double)xin[r] * (double)acc[h];
        } else {
          const double y = b[r] - (double)acc[h];
          yout[r] = (VT)y;
          dot += y * y;
        }
      }
    }
  }
  return dot;
}

struct BatchArgs {
  int nsys;
  const int32_t* sys_n;
  const int64_t* sys_ip_off;
  const int64_t* sys_nnz_off;
  const int64_t* sys_vec_off;
  const int32_t* order;
  const int32_t* indptr;
  const int32_t* indices;
  const double* vals;
  const float* vals32;
  const double* B;
  double* X;
  double tol;
  int maxiter;
  int precond;
  double inner_tol;
  int max_outer;
  int* counter;
  cc_pcg_result* results;
};

template <int NT, int T, int RPT, bool MIXED>
__global__ __launch_bounds__(NT) void k_batched_pcg(BatchArgs a) {
  constexpr int NW = NT / 64;
  extern __shared__ double lds[];
  __shared__ double red[2 * NW];
  __shared__ int s_sys;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_sys = atomicAdd(a.counter, 1);
    __syncthreads();
    const int w = s_sys;
    if (w >= a.nsys) return;
    const int sid = a.order[w];
    const int n = a.sys_n[sid];
    if (n < 0 || n > RPT * NT) {
      // The caller's n_max was below this system's size: the launch shape
      // holds only RPT * NT rows per system in registers and LDS. Report
      // the system as a bad argument and leave its solution untouched.
      if (threadIdx.x == 0) {
        cc_pcg_result res;
        res.iterations = 0;
        res.status = CC_E_ARG;
        res.outer = 0;
        res.inner_total = 0;
        res.recursive_rel = 0.0;
        res.true_rel = 0.0;
        a.results[sid] = res;
      }
      continue;
    }
    const int* ip = a.indptr + a.sys_ip_off[sid];
    const int* ix = a.indices + a.sys_nnz_off[sid];
    const double* va = a.vals + a.sys_nnz_off[sid];
    const double* b = a.B + a.sys_vec_off[sid];
    double* xo = a.X + a.sys_vec_off[sid];

    double d[RPT], bb_[RPT];
    double bsq = 0.0;
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
      const int i = threadIdx.x + j * NT;
      d[j] = 0.0;
      bb_[j] = 0.0;
      if (i < n) {
        double s = 0.0;
        for (int e = ip[i]; e < ip[i + 1]; ++e)
          if (ix[e] == i) s += va[e];
        d[j] = s;
        bb_[j] = b[i];
        bsq += bb_[j] * bb_[j];
      }
    }
    bsq = wg_sum<NW>(bsq, red);
    double bnorm = sqrt(bsq);
    if (bnorm == 0.0) bnorm = 1.0;

    int status = CC_MAXITER, iters = 0, outer = 1, inner_total = 0;
    double rel = sqrt(bsq) / bnorm;

    if constexpr (!MIXED) {
      double* P = lds;
      double* Q = lds + (size_t)RPT * NT;
      double x[RPT], r[RPT], z[RPT];
      double rho = 0.0;
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int i = threadIdx.x + j * NT;
        x[j] = 0.0;
        r[j] = bb_[j];
        z[j] = (a.precond && fabs(d[j]) > 0.0) ? r[j] / d[j] : r[j];
        if (i < n) P[i] = z[j];
        rho += r[j] * z[j];
      }
      rho = wg_sum<NW>(rho, red);  // also publishes P
      for (int k = 1; k <= a.maxiter; ++k) {
        double pq = team_spmv<double, double, T, 0>(n, ip, ix, va, P, Q, nullptr);
        pq = wg_sum<NW>(pq, red);  // also publishes Q
        if (!(pq > 0.0 && isfinite(pq))) {
          status = CC_BREAKDOWN;
          iters = k;
          break;
        }
        const double alpha = rho / pq;
        double rr = 0.0, rz = 0.0;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int i = threadIdx.x + j * NT;
          if (i < n) {
            x[j] += alpha * P[i];
            r[j] -= alpha * Q[i];
            z[j] = (a.precond && fabs(d[j]) > 0.0) ? r[j] / d[j] : r[j];
            rr += r[j] * r[j];
            rz += r[j] * z[j];
          }
        }
        wg_sum2<NW>(rr, rz, red);
        rel = sqrt(rr) / bnorm;
        iters = k;
        if (rel <= a.tol) {
          status = CC_CONVERGED;
          break;
        }
        if (k == a.maxiter) break;
        const double beta = rz / rho;
        rho = rz;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int i = threadIdx.x + j * NT;
          if (i < n) P[i] = z[j] + beta * P[i];
        }
        __syncthreads();
      }
      // Float64 true residual ||b - A x|| with x staged in LDS.
      __syncthreads();
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int i = threadIdx.x + j * NT;
        if (i < n) {
          P[i] = x[j];
          xo[i] = x[j];
        }
      }
      __syncthreads();
      double tr = team_spmv<double, double, T, 1>(n, ip, ix, va, P, Q, b);
      tr = wg_sum<NW>(tr, red);
      inner_total = iters;
      if (threadIdx.x == 0) {
        cc_pcg_result res;
        res.iterations = iters;
        res.status = status;
        res.outer = 1;
        res.inner_total = iters;
        res.recursive_rel = rel;
        res.true_rel = sqrt(tr) / bnorm;
        a.results[sid] = res;
      }
    } else {
      // Outer float64 state.
      const float* va32 = a.vals32 + a.sys_nnz_off[sid];
      double x64[RPT], r64[RPT];
      double rr64 = 0.0;
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        x64[j] = 0.0;
        r64[j] = bb_[j];
        rr64 += r64[j] * r64[j];
      }
      rr64 = wg_sum<NW>(rr64, red);
      double orel = sqrt(rr64) / bnorm;
      float* P = reinterpret_cast<float*>(lds);
      float* Q = P + (size_t)RPT * NT;
      double* X64 = lds;  // reused for the float64 residual product
      double* R64 = lds + (size_t)RPT * NT;
      int sweeps = 0;
      status = CC_MAXITER;
      while (true) {
        if (orel <= a.tol) {
          status = CC_CONVERGED;
          break;
        }
        if (sweeps >= a.max_outer || inner_total >= a.maxiter) break;
        ++sweeps;
        const double scale = sqrt(rr64);
        const double inv = scale > 0.0 ? 1.0 / scale : 0.0;
        float x[RPT], r[RPT], z[RPT];
        double rho = 0.0, rin = 0.0;
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int i = threadIdx.x + j * NT;
          x[j] = 0.f;
          r[j] = (float)(r64[j] * inv);
          z[j] = (a.precond && fabs(d[j]) > 0.0) ? (float)((double)r[j] / d[j]) : r[j];
          if (i < n) P[i] = z[j];
          rho += (double)r[j] * (double)z[j];
          rin += (double)r[j] * (double)r[j];
        }
        wg_sum2<NW>(rho, rin, red);
        double ibn = sqrt(rin);
        if (ibn == 0.0) ibn = 1.0;
        const int budget = a.maxiter - inner_total;
        int k = 0;
        for (k = 1; k <= budget; ++k) {
          double pq = team_spmv<float, float, T, 0>(n, ip, ix, va32, P, Q, nullptr);
          pq = wg_sum<NW>(pq, red);
          if (!(pq > 0.0 && isfinite(pq))) break;
          const double alpha = rho / pq;
          double rr = 0.0, rz = 0.0;
#pragma unroll
          for (int j = 0; j < RPT; ++j) {
            const int i = threadIdx.x + j * NT;
            if (i < n) {
              x[j] = (float)((double)x[j] + alpha * (double)P[i]);
              r[j] = (float)((double)r[j] - alpha * (double)Q[i]);
              z[j] = (a.precond && fabs(d[j]) > 0.0) ? (float)((double)r[j] / d[j]) : r[j];
              rr += (double)r[j] * (double)r[j];
              rz += (double)r[j] * (double)z[j];
            }
          }
          wg_sum2<NW>(rr, rz, red);
          if (sqrt(rr) / ibn <= a.inner_tol) break;
          const double beta = rz / rho;
          rho = rz;
#pragma unroll
          for (int j = 0; j < RPT; ++j) {
            const int i = threadIdx.x + j * NT;
            if (i < n) P[i] = (float)((double)z[j] + beta * (double)P[i]);
          }
          __syncthreads();
        }
        inner_total += (k > budget ? budget : k);
        // x64 += scale * x; r64 = b - A64 x64 in float64.
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int i = threadIdx.x + j * NT;
          x64[j] += scale * (double)x[j];
          if (i < n) X64[i] = x64[j];
        }
        __syncthreads();
        double t = team_spmv<double, double, T, 1>(n, ip, ix, va, X64, R64, b);
        rr64 = wg_sum<NW>(t, red);
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int i = threadIdx.x + j * NT;
          r64[j] = (i < n) ? R64[i] : 0.0;
        }
        orel = sqrt(rr64) / bnorm;
      }
#pragma unroll
      for (int j = 0; j < RPT; ++j) {
        const int i = threadIdx.x + j * NT;
        if (i < n) xo[i] = x64[j];
      }
      outer = sweeps;
      if (threadIdx.x == 0) {
        cc_pcg_result res;
        res.iterations = inner_total;
        res.status = status;
        res.outer = outer;
        res.inner_total = inner_total;
        res.recursive_rel = orel;
        res.true_rel = orel;
        a.results[sid] = res;
      }
    }
  }
}

__global__ void k_reset_counter(int* c) { *c = 0; }

template <int NT, int T, int RPT, bool MIXED>
static int launch_batched(const BatchArgs& a, int grid, hipStream_t s) {
  const size_t lds = 2 * (size_t)RPT * NT * sizeof(double);
  auto kern = k_batched_pcg<NT, T, RPT, MIXED>;
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(NT), lds, s, a);
  return (int)hipGetLastError();
}

template <int NT, int T, bool MIXED>
static int dispatch_rpt(const BatchArgs& a, int nmax, int grid, hipStream_t s) {
  const int rpt = (nmax + NT - 1) / NT;
  if (rpt <= 1) return launch_batched<NT, T, 1, MIXED>(a, grid, s);
  if (rpt <= 2) return launch_batched<NT, T, 2, MIXED>(a, grid, s);
  if (rpt <= 4) return launch_batched<NT, T, 4, MIXED>(a, grid, s);
  if (rpt <= 8) return launch_batched<NT, T, 8, MIXED>(a, grid, s);
  if (rpt <= 16 && NT * 16 <= 8192) return launch_batched<NT, T, (NT * 16 <= 8192 ? 16 : 8), MIXED>(a, grid, s);
  return CC_E_ARG;
}

template <int NT, bool MIXED>
static int dispatch_team(const BatchArgs& a, int team, int nmax, int grid, hipStream_t s) {
  switch (team) {
    case 4: return dispatch_rpt<NT, 4, MIXED>(a, nmax, grid, s);
    case 8: return dispatch_rpt<NT, 8, MIXED>(a, nmax, grid, s);
    case 16: return dispatch_rpt<NT, 16, MIXED>(a, nmax, grid, s);
    case 32: return dispatch_rpt<NT, 32, MIXED>(a, nmax, grid, s);
    case 64: return dispatch_rpt<NT, 64, MIXED>(a, nmax, grid, s);
    default: return CC_E_ARG;
  }
}

template <bool MIXED>
static int dispatch_block(const BatchArgs& a, int block, int team, int nmax, int grid, hipStream_t s) {
  switch (block) {
    case 256: return dispatch_team<256, MIXED>(a, team, nmax, grid, s);
    case 512: return dispatch_team<512, MIXED>(a, team, nmax, grid, s);
    case 1024: return dispatch_team<1024, MIXED>(a, team, nmax, grid, s);
    default: return CC_E_ARG;
  }
}

}  // namespace cc

using namespace cc;

extern "C" int cc_pcg_batched(int32_t nsys, int32_t n_max, const int32_t* sys_n, const int64_t* sys_indptr_off,
                              const int64_t* sys_nnz_off, const int64_t* sys_vec_off, const int32_t* order,
                              const int32_t* indptr, const int32_t* indices, const double* vals,
                              const float* vals32, const double* B, double* X, double tol, int32_t maxiter,
                              int32_t preconditioner, int32_t precision, double inner_tol, int32_t max_outer,
                              int32_t block_threads, int32_t team_lanes, int32_t grid, int32_t* counter,
                              cc_pcg_result* results, cc_stream_t stream) {
  if (nsys <= 0 || n_max <= 0 || n_max > 8192 || !sys_n || !sys_indptr_off || !sys_nnz_off || !sys_vec_off ||
      !order || !indptr || !indices || !vals || !B || !X || !counter || !results || maxiter < 0)
    return CC_E_ARG;
  if (precision == 1 && !vals32) return CC_E_ARG;
  if (grid <= 0) grid = nsys;
  hipStream_t s = (hipStream_t)stream;
  BatchArgs a{nsys, sys_n, sys_indptr_off, sys_nnz_off, sys_vec_off, order, indptr, indices, vals, vals32, B, X,
              tol, maxiter, preconditioner, inner_tol, max_outer, counter, results};
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s, counter);
  CC_CHECK(hipGetLastError());
  prof_begin_batched(s);
  int rc = precision == 1 ? dispatch_block<true>(a, block_threads, team_lanes, n_max, grid, s)
                          : dispatch_block<false>(a, block_threads, team_lanes, n_max, grid, s);
  prof_end_batched(s);
  return rc;
}
