// pcg.hip - CSR SpMV/SpMM and device-resident Jacobi-PCG for gfx950.
//
// One HIP stream, no host synchronisation inside an iteration: every scalar of
// the CG recurrence (alpha, beta, ||r||, the stopping and breakdown decisions)
// lives on the device. One iteration is five launches:
//   S   q = A p                       (+ per-block partials of p.q)
//   F1  alpha = rho / p.q, breakdown  (one block)
//   U   x += a p, r -= a q, z = r/d   (+ per-block partials of r.r, r.z)
//   F2  rel, stop test, beta, rho     (one block)
//   P   p = z + b p
// Cross-block sums never use atomics: the one-block finalize kernels reduce
// the producers' per-block partials in a fixed order, so results are bitwise
// reproducible run to run, and the streaming kernels read only the k
// per-column scalars. The host polls the per-column status every
// `check_every` iterations; kernels launched after every column has stopped
// return at once.
//
// The recurrence and the stopping / breakdown rules follow the reference GPU
// loop, code/fast_solver_backends.py:233-274 (x0 = 0, r = b, z = r/d with
// zero diagonal -> identity, rel = ||r||/||b|| <= tol non-strict, breakdown
// when p.Ap <= 0 or non-finite, iterations = k at exit).
//
// SpMM layout: a row team of T = K*NZ lanes; lane (z, c) takes nonzeros z,
// z+NZ, ... of the row for right-hand side c (X and Y row-major [n][K], so the
// K values of a gathered row are one contiguous segment). Each block owns a
// contiguous chunk of rows, and chunks are dealt to blocks so that the blocks
// the hardware places on one XCD (block id mod 8) cover adjacent chunks: the
// x rows a chunk gathers then mostly hit that XCD's L2.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <vector>

#include "../../include/cutcell_hip.h"
#include "common.h"
#include "spmm_util.h"

namespace cc {


// ---------------------------------------------------------------------------
// Optional per-kernel-class event timing (process-wide, not thread-safe).
// ---------------------------------------------------------------------------

struct Profiler {
  bool on = false;
  uint32_t classes = 0;  // bit c: class c is timed
  bool open = false;     // a timed class's first mark is down (pairs never nest)
  int armed = -1;        // class whose single launch takes the next event pair
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<int, size_t>> pending;
  double ms[kProfClasses] = {0};
  long long count[kProfClasses] = {0};
};
static Profiler g_prof;

static void prof_mark(hipStream_t s) {
  if (g_prof.used >= g_prof.pool.size()) {
    // No system-scope fence at the marks: they only order work on one
    // stream, and a fence would write back L2 between every two kernels.
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return;
    g_prof.pool.push_back(e);
  }
  hipEventRecord(g_prof.pool[g_prof.used], s);
  ++g_prof.used;
}

void prof_begin(int cls, hipStream_t s) {
  if (!g_prof.on || !(g_prof.classes >> cls & 1u)) return;
  g_prof.pending.push_back({cls, g_prof.used});
  prof_mark(s);
  g_prof.open = true;
}

void prof_end(hipStream_t s) {
  if (!g_prof.on || !g_prof.open) return;
  prof_mark(s);
  g_prof.open = false;
}

void prof_arm(int cls) {
  if (!g_prof.on || !(g_prof.classes >> cls & 1u)) return;
  g_prof.armed = cls;
}

void prof_disarm() { g_prof.armed = -1; }

void prof_take(hipEvent_t* start, hipEvent_t* stop) {
  *start = *stop = nullptr;
  if (g_prof.armed < 0) return;
  while (g_prof.pool.size() < g_prof.used + 2) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return;
    g_prof.pool.push_back(e);
  }
  g_prof.pending.push_back({g_prof.armed, g_prof.used});
  *start = g_prof.pool[g_prof.used];
  *stop = g_prof.pool[g_prof.used + 1];
  g_prof.used += 2;
  g_prof.armed = -1;
}

// Position in the list of timed intervals not yet collected.
static size_t prof_pending_mark() { return g_prof.pending.size(); }

// Of the intervals of class cls recorded since mark, keep the first `live`
// and drop the rest (launches queued after every column had stopped, whose
// kernels exit at once).
static void prof_keep(int cls, size_t mark, long long live) {
  long long seen = 0;
  for (size_t i = mark; i < g_prof.pending.size(); ++i)
    if (g_prof.pending[i].first == cls && seen++ >= live) g_prof.pending[i].first = -1;
}

// Call only after the stream has been synchronised.
void prof_collect() {
  if (!g_prof.on) return;
  for (auto& pr : g_prof.pending) {
    if (pr.first < 0) continue;
    float t = 0.f;
    if (pr.second + 1 < g_prof.used &&
        hipEventElapsedTime(&t, g_prof.pool[pr.second], g_prof.pool[pr.second + 1]) == hipSuccess) {
      g_prof.ms[pr.first] += t;
      g_prof.count[pr.first] += 1;
    }
  }
  g_prof.pending.clear();
  g_prof.used = 0;
}

// ---------------------------------------------------------------------------
// SpMM launch configuration (per vector width and k; set by the autotuner or
// the defaults). Index [0] holds float64 vectors, [1] float32 ones.
// ---------------------------------------------------------------------------

struct SpmmConfig {
  int nz;      // nonzero slots per row team ((K/cpl)*nz <= 64); 0 = pick from the row length
  int grid;    // blocks; 0 = min(kMaxParts, rows / teams per block)
  int unroll;  // 1, 4 or 8 (software-pipelined) loads in flight per lane; 0 = 4;
               // kUnrollWindow: the nonzero-window kernel (k = 1 only)
  int cpl;     // right-hand sides per lane (vector width); 0 = widest 16-byte load
  int stream;  // 1: nontemporal matrix loads (not with unroll 1)
  int nsub;    // sweep steps: chunks per block, taken grid-strided; 0 = 1
};
constexpr int kUnrollWindow = 16;
constexpr int kMaxSweep = 64;
static SpmmConfig g_spmm_cfg[2][17] = {};

template <typename VT>
static inline const SpmmConfig& spmm_cfg(int k) {
  return g_spmm_cfg[sizeof(VT) == 4 ? 1 : 0][k];
}

// ---------------------------------------------------------------------------
// Reductions.
// ---------------------------------------------------------------------------

// Sum a thread's value over all threads of the block that share its column
// c = threadIdx.x % K; thread c < K writes the block total to out[c].
template <int K>
__device__ inline void block_col_sum(double v, double* out, double* smem) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int o = K; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
  if (lane < K) smem[wave * K + lane] = v;
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += smem[w * K + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// Reduce partial[nparts][PAIRS] over the parts, in a fixed order, into
// out[PAIRS] (shared memory). PAIRS must divide blockDim.x. Each thread keeps
// four independent running sums (parts sub, sub + nsub, sub + 2 nsub,
// sub + 3 nsub, then the next four ...) so four loads are in flight, and the
// sums are combined in a fixed order: the result is bitwise reproducible.
template <int PAIRS>
__device__ inline void reduce_partials(const double* __restrict__ partial, int nparts, double* out, double* smem) {
  const int pair = threadIdx.x % PAIRS;
  const int sub = threadIdx.x / PAIRS;
  const int nsub = blockDim.x / PAIRS;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int b = sub;
  for (; b + 3 * nsub < nparts; b += 4 * nsub) {
    s0 += partial[(size_t)b * PAIRS + pair];
    s1 += partial[(size_t)(b + nsub) * PAIRS + pair];
    s2 += partial[(size_t)(b + 2 * nsub) * PAIRS + pair];
    s3 += partial[(size_t)(b + 3 * nsub) * PAIRS + pair];
  }
  for (; b < nparts; b += nsub) s0 += partial[(size_t)b * PAIRS + pair];
  smem[threadIdx.x] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (threadIdx.x < PAIRS) {
    double t = 0.0;
    for (int j = 0; j < nsub; ++j) t += smem[j * PAIRS + threadIdx.x];
    out[threadIdx.x] = t;
  }
  __syncthreads();
}

// Reduce partial[nparts][PAIRS] (nparts <= kMaxParts) into out[PAIRS] with a
// block of NT threads: every thread issues all of its kMaxParts * PAIRS / NT
// loads at once (one round of memory latency), sums them pairwise, and the
// block combines the per-thread sums by a tree in LDS. The order is fixed, so
// the result is bitwise reproducible. smem holds NT doubles.
template <int PAIRS, int NT>
__device__ inline void reduce_partials_wide(const double* __restrict__ partial, int nparts, double* out,
                                            double* smem) {
  static_assert(NT % PAIRS == 0 && (kMaxParts * PAIRS) % NT == 0, "shape");
  constexpr int PER = kMaxParts * PAIRS / NT;
  const int total = nparts * PAIRS;
  double v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int e = threadIdx.x + j * NT;
    v[j] = e < total ? partial[e] : 0.0;
  }
#pragma unroll
  for (int w = 1; w < PER; w <<= 1)
#pragma unroll
    for (int j = 0; j + w < PER; j += 2 * w) v[j] += v[j + w];
  // Element e = t + j NT belongs to pair e % PAIRS = t % PAIRS.
  smem[threadIdx.x] = v[0];
  __syncthreads();
  for (int w = NT / 2; w >= PAIRS; w >>= 1) {
    if (threadIdx.x < w) smem[threadIdx.x] += smem[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < PAIRS) out[threadIdx.x] = smem[threadIdx.x];
  __syncthreads();
}

constexpr int kFinThreads = 1024;

// alpha = rho / p.q; breakdown when p.q <= 0 or non-finite.
template <int K>
__device__ inline void finalize_alpha(const double* __restrict__ partial, int nparts, ColState* st, int iter) {
  __shared__ double smem[kFinThreads];
  __shared__ double den[K];
  reduce_partials_wide<K, kFinThreads>(partial, nparts, den, smem);
  if (threadIdx.x < K) {
    const int c = threadIdx.x;
    if (st[c].status == CC_RUNNING) {
      const double dn = den[c];
      if (!(dn > 0.0 && isfinite(dn))) {
        st[c].status = CC_BREAKDOWN;
        st[c].iterations = iter;
        st[c].alpha = 0.0;
      } else {
        st[c].alpha = st[c].rho / dn;
      }
    }
  }
}

// rel = ||r||/||b||; stop when rel <= tol (or iter == maxiter); otherwise
// beta = (r.z)_new / rho and rho = (r.z)_new.
template <int K>
__device__ inline void finalize_beta(const double* __restrict__ partial, int nparts, ColState* st, int iter,
                                     double tol, int maxiter, double* hist) {
  __shared__ double smem[kFinThreads];
  __shared__ double sums[2 * K];
  reduce_partials_wide<2 * K, kFinThreads>(partial, nparts, sums, smem);
  if (threadIdx.x < K) {
    const int c = threadIdx.x;
    if (st[c].status == CC_RUNNING) {
      const double rr = sums[c];
      const double rz = sums[K + c];
      const double rel = sqrt(rr) / st[c].bnorm;
      st[c].rel = rel;
      st[c].iterations = iter;
      if (hist) hist[(size_t)iter * K + c] = rel;
      if (rel <= tol) {
        st[c].status = CC_CONVERGED;
      } else if (iter >= maxiter) {
        st[c].status = CC_MAXITER;
      } else {
        st[c].beta = rz / st[c].rho;
        st[c].rho = rz;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_diag(int n, const int* __restrict__ indptr, const int* __restrict__ indices,
                                              const double* __restrict__ vals, double* __restrict__ diag) {
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  double d = 0.0;
  for (int j = indptr[row]; j < indptr[row + 1]; ++j)
    if (indices[j] == row) d += vals[j];
  diag[row] = d;
}

// MODE 0: Y = A X, partial[block][c] = sum_rows X.Y      (p.q of CG)
// MODE 1: Y = B - A X, partial[block][c] = sum_rows Y.Y  (true residual)
// A row team has KL = K/CPL lanes per nonzero slot (each lane owns CPL
// consecutive right-hand sides, loaded as one vector) and NZ slots.
// NTL: load the matrix (read once per product) with nontemporal loads so it
// does not evict the gathered X rows from L2; chosen by the autotuner.
// nsub: the rows are cut into gridDim.x * nsub chunks of equal nonzero
// counts; block b takes chunks b, b + G, b + 2G, ... in that order, so at any
// time the whole grid streams one compact stretch of the matrix (DRAM pages
// stay open) instead of G stretches far apart.
template <typename MT, typename VT, int K, int NZ, int MODE, int UNR, int CPL, bool NTL>
__global__ __launch_bounds__(256) void k_spmm(int n, const int* __restrict__ indptr, const int* __restrict__ indices,
                                              const MT* __restrict__ vals, const VT* __restrict__ X,
                                              VT* __restrict__ Y, const double* __restrict__ B,
                                              double* __restrict__ partial, const ColState* __restrict__ st,
                                              int nsub) {
  constexpr int KL = K / CPL;
  constexpr int T = KL * NZ;
  constexpr int TPW = 64 / T;
  __shared__ double smem[4 * 16];
  if (all_stopped<K>(st)) return;
  using AT = typename Accum<VT>::type;
  const int lane = threadIdx.x & 63;
  const int c0 = (lane % KL) * CPL;
  const int zs = (lane / KL) % NZ;
  const int wave = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int tpb = wpb * TPW;
  const int team = wave * TPW + lane / T;
  // XCD-aware chunks of rows: chunks hold equal shares of the nonzeros (rows
  // along a cut are up to three times longer than the rest), and within one
  // sweep step the blocks one XCD runs (block id mod 8) take adjacent chunks.
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int chunk0 = (G % kXcds == 0) ? (b % kXcds) * (G / kXcds) + b / kXcds : b;
  const int NC = G * nsub;
  __shared__ int s_rows[2][2];
  double dot[CPL];
#pragma unroll
  for (int m = 0; m < CPL; ++m) dot[m] = 0.0;
  for (int sub = 0; sub < nsub; ++sub) {
  const int chunk = sub * G + chunk0;
  if (wave == 0) {
    const long long base = indptr[0];
    const long long total = (long long)indptr[n] - base;
    const int a0 = chunk == 0 ? 0 : row_lower_bound(indptr, n, base + total * chunk / NC);
    const int a1 = chunk == NC - 1 ? n : row_lower_bound(indptr, n, base + total * (chunk + 1) / NC);
    if (lane == 0) {
      s_rows[sub & 1][0] = a0;
      s_rows[sub & 1][1] = a1;
    }
  }
  // Double-buffered: the next step's bounds go to the other slot, which every
  // wave finished reading before this barrier.
  __syncthreads();
  const int r0 = s_rows[sub & 1][0];
  const int r1 = s_rows[sub & 1][1];
  // UNR == 8 carries state from one row of the team to the next: the next
  // row's bounds (loaded when the current row starts) and its first chunk of
  // indices and values (loaded during the current row's last chunk), so a row
  // never starts with a cold chain of row pointer, entries and gathers.
  constexpr int U = 8;
  int qa[U];
  AT va[U];
  int nbeg = 0, nlen = 0;
  // The next 8 entries of this lane at j0 of a row of length len (entries past
  // the end read X row `dummy`, always in range, and carry a zero value).
  auto load_chunk = [&](const int* ip, const MT* vp, int j0, int len_, int dummy, int* q, AT* v) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int jj = j0 + u * NZ;
      q[u] = jj < len_ ? stream_ld<NTL>(ip + jj) : dummy;
      v[u] = jj < len_ ? (AT)stream_ld<NTL>(vp + jj) : (AT)0;
    }
  };
  if constexpr (UNR == 8) {
    const int row = r0 + team;
    if (row < r1) {
      nbeg = indptr[row];
      nlen = indptr[row + 1] - nbeg;
      load_chunk(indices + nbeg, vals + nbeg, zs, nlen, row, qa, va);
    }
  }
  for (int row = r0 + team; row < r1; row += tpb) {
    int beg, len;
    if constexpr (UNR == 8) {
      beg = nbeg;
      len = nlen;
    } else {
      beg = indptr[row];
      len = indptr[row + 1] - beg;
    }
    const int* __restrict__ ixp = indices + beg;
    const MT* __restrict__ vap = vals + beg;
    AT acc[CPL];
#pragma unroll
    for (int m = 0; m < CPL; ++m) acc[m] = 0;
    int j = zs;
    if constexpr (UNR == 4) {
      AT a1[CPL], a2[CPL], a3[CPL];
#pragma unroll
      for (int m = 0; m < CPL; ++m) a1[m] = a2[m] = a3[m] = 0;
      for (; j + 3 * NZ < len; j += 4 * NZ) {
        const int q0 = stream_ld<NTL>(ixp + j), q1 = stream_ld<NTL>(ixp + j + NZ);
        const int q2 = stream_ld<NTL>(ixp + j + 2 * NZ), q3 = stream_ld<NTL>(ixp + j + 3 * NZ);
        const AT v0 = (AT)stream_ld<NTL>(vap + j), v1 = (AT)stream_ld<NTL>(vap + j + NZ);
        const AT v2 = (AT)stream_ld<NTL>(vap + j + 2 * NZ), v3 = (AT)stream_ld<NTL>(vap + j + 3 * NZ);
        VT x0[CPL], x1[CPL], x2[CPL], x3[CPL];
        vload<VT, CPL>(X + (size_t)q0 * K + c0, x0);
        vload<VT, CPL>(X + (size_t)q1 * K + c0, x1);
        vload<VT, CPL>(X + (size_t)q2 * K + c0, x2);
        vload<VT, CPL>(X + (size_t)q3 * K + c0, x3);
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
          acc[m] += v0 * (AT)x0[m];
          a1[m] += v1 * (AT)x1[m];
          a2[m] += v2 * (AT)x2[m];
          a3[m] += v3 * (AT)x3[m];
        }
      }
#pragma unroll
      for (int m = 0; m < CPL; ++m) acc[m] = (acc[m] + a1[m]) + (a2[m] + a3[m]);
    }
    if constexpr (UNR == 8) {
      // Software-pipelined: the indices and values of the next 8 entries of
      // this lane (or, at the row's last chunk, the next row's first 8) are
      // in flight while the X rows of the current 8 are gathered, so each
      // step waits on one memory latency (the gather).
      const int nrow = row + tpb;
      const bool has_next = nrow < r1;
      nbeg = has_next ? indptr[nrow] : 0;
      nlen = has_next ? indptr[nrow + 1] - nbeg : 0;
      const int* __restrict__ nixp = indices + nbeg;
      const MT* __restrict__ nvap = vals + nbeg;
      int qb[U];
      AT vb[U];
      AT a1[CPL];
#pragma unroll
      for (int m = 0; m < CPL; ++m) a1[m] = 0;
      if (zs >= len) {
        load_chunk(nixp, nvap, zs, nlen, row, qa, va);
      } else {
        for (int j0 = zs; j0 < len; j0 += U * NZ) {
          const int jn = j0 + U * NZ;
          if (jn < len)
            load_chunk(ixp, vap, jn, len, row, qb, vb);
          else
            load_chunk(nixp, nvap, zs, nlen, row, qb, vb);
          VT xv[U][CPL];
#pragma unroll
          for (int u = 0; u < U; ++u) vload<VT, CPL>(X + (size_t)qa[u] * K + c0, xv[u]);
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const bool live = j0 + u * NZ < len;
#pragma unroll
            for (int m = 0; m < CPL; ++m) {
              const AT t = live ? va[u] * (AT)xv[u][m] : (AT)0;
              if (u & 1)
                a1[m] += t;
              else
                acc[m] += t;
            }
          }
#pragma unroll
          for (int u = 0; u < U; ++u) {
            qa[u] = qb[u];
            va[u] = vb[u];
          }
        }
      }
#pragma unroll
      for (int m = 0; m < CPL; ++m) acc[m] += a1[m];
      j = len;
    }
    for (; j < len; j += NZ) {
      const AT v = (AT)stream_ld<NTL>(vap + j);
      VT xv[CPL];
      vload<VT, CPL>(X + (size_t)stream_ld<NTL>(ixp + j) * K + c0, xv);
#pragma unroll
      for (int m = 0; m < CPL; ++m) acc[m] += v * (AT)xv[m];
    }
#pragma unroll
    for (int o = KL; o < T; o <<= 1)
#pragma unroll
      for (int m = 0; m < CPL; ++m) acc[m] += __shfl_xor(acc[m], o, 64);
    if (zs == 0) {
      const size_t e = (size_t)row * K + c0;
      VT yv[CPL];
      if (MODE == 0) {
#pragma unroll
        for (int m = 0; m < CPL; ++m) yv[m] = (VT)acc[m];
        vstore<VT, CPL>(Y + e, yv);
        if (partial) {
          VT xv[CPL];
          vload<VT, CPL>(X + e, xv);
#pragma unroll
          for (int m = 0; m < CPL; ++m) dot[m] += (double)xv[m] * (double)yv[m];
        }
      } else {
#pragma unroll
        for (int m = 0; m < CPL; ++m) {
          const double y = B[e + m] - (double)acc[m];
          yv[m] = (VT)y;
          dot[m] += y * y;
        }
        vstore<VT, CPL>(Y + e, yv);
      }
    }
  }
  }  // sweep step
  if (partial) {
    // Sum over the lanes that own the same columns, then over waves.
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

// ---------------------------------------------------------------------------
// Window kernel for one right-hand side (k = 1). A block streams the nonzeros
// of its row range (an equal share of the nonzeros, as in k_spmm) in windows
// of kWin entries: every lane loads 4 consecutive entries per step with
// 16-byte loads, whatever the row lengths, gathers their X values and writes
// the products to LDS in nonzero order. Then 16-lane teams sum each row's
// segment of the window in a fixed order (lane-strided, then a shuffle tree).
// A row that crosses the window's end hands its partial sum to the next
// window, so a row of any length is summed without atomics. The next
// window's indices and values are loaded before the row sums of the current
// one, so the matrix stream stays in flight through the sum phase.
// ---------------------------------------------------------------------------

typedef int ivec4 __attribute__((ext_vector_type(4)));
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));

constexpr int kWinBlock = 512;                  // threads per block
constexpr int kWinReps = 2;                     // 4-entry groups per lane and window
constexpr int kWin = kWinBlock * 4 * kWinReps;  // 4096 nonzeros
constexpr int kRowLanes = 16;                   // lanes per row team in the sum phase
constexpr int kWinTeams = kWinBlock / kRowLanes;

// Four consecutive matrix values (16-byte aligned).
template <bool NTL>
__device__ inline void load_vals4(const double* p, double* v) {
  const dvec2 a = stream_ld<NTL>(reinterpret_cast<const dvec2*>(p));
  const dvec2 c = stream_ld<NTL>(reinterpret_cast<const dvec2*>(p) + 1);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = c.x;
  v[3] = c.y;
}
template <bool NTL>
__device__ inline void load_vals4(const float* p, float* v) {
  const fvec4 a = stream_ld<NTL>(reinterpret_cast<const fvec4*>(p));
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
}

// One window's entries held by a lane: indices (-1 = no entry) and values.
template <typename MT>
struct WinRegs {
  ivec4 q[kWinReps];
  MT v[kWinReps][4];
};

// Load the lane's entries of the window [s, e) (s a multiple of 4); entries
// before nb (the block's first) or at or past e are not loaded.
template <typename MT, bool NTL>
__device__ inline void win_load(WinRegs<MT>& w, const int* __restrict__ indices, const MT* __restrict__ vals, int s,
                                int nb, int e, int t) {
#pragma unroll
  for (int rep = 0; rep < kWinReps; ++rep) {
    const int j = s + (rep * kWinBlock + t) * 4;
    if (s + rep * kWinBlock * 4 >= e) {
      // A step past the window's end (uniform over the block).
      w.q[rep] = ivec4{-1, -1, -1, -1};
#pragma unroll
      for (int u = 0; u < 4; ++u) w.v[rep][u] = (MT)0;
    } else if (j >= nb && j + 4 <= e) {
      w.q[rep] = stream_ld<NTL>(reinterpret_cast<const ivec4*>(indices + j));
      load_vals4<NTL>(vals + j, w.v[rep]);
    } else {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int jj = j + u;
        const bool ok = jj >= nb && jj < e;
        w.q[rep][u] = ok ? indices[jj] : -1;
        w.v[rep][u] = ok ? vals[jj] : (MT)0;
      }
    }
  }
}

template <typename MT, typename VT, int MODE, bool NTL>
__global__ __launch_bounds__(kWinBlock) void k_spmv_win(int n, const int* __restrict__ indptr,
                                                        const int* __restrict__ indices, const MT* __restrict__ vals,
                                                        const VT* __restrict__ X, VT* __restrict__ Y,
                                                        const double* __restrict__ B, double* __restrict__ partial,
                                                        const ColState* __restrict__ st, int nsub) {
  using AT = typename Accum<VT>::type;
  __shared__ __align__(16) AT prod[kWin];
  __shared__ double smem[kWinBlock / 64];
  __shared__ int s_rows[2][2];
  __shared__ int s_next[2];  // first unfinished row, the row s_carry belongs to (-1: none)
  __shared__ AT s_carry;
  if (all_stopped<1>(st)) return;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int G = gridDim.x;
  const int b = blockIdx.x;
  const int chunk0 = (G % kXcds == 0) ? (b % kXcds) * (G / kXcds) + b / kXcds : b;
  const int NC = G * nsub;
  const int team = t / kRowLanes;
  const int tl = t % kRowLanes;
  double dot = 0.0;
  // Sweep steps as in k_spmm: chunk sub * G + chunk0 at step sub.
  for (int sub = 0; sub < nsub; ++sub) {
    const int chunk = sub * G + chunk0;
    if (wave == 0) {
      const long long base = indptr[0];
      const long long total = (long long)indptr[n] - base;
      const int a0 = chunk == 0 ? 0 : row_lower_bound(indptr, n, base + total * chunk / NC);
      const int a1 = chunk == NC - 1 ? n : row_lower_bound(indptr, n, base + total * (chunk + 1) / NC);
      if (lane == 0) {
        s_rows[sub & 1][0] = a0;
        s_rows[sub & 1][1] = a1;
      }
    }
    __syncthreads();
    const int r0 = s_rows[sub & 1][0];
    const int r1 = s_rows[sub & 1][1];
    if (r0 < r1) {
      const int nb = indptr[r0];
      const int ne = indptr[r1];
      int rc = r0;          // first row not yet finished
      int carry_row = -1;   // row whose partial sum carry holds
      AT carry = 0;
      int s = nb & ~3;      // window start, a multiple of 4 (16-byte loads)
      int e = (ne - s > kWin) ? s + kWin : ne;
      WinRegs<MT> w;
      win_load<MT, NTL>(w, indices, vals, s, nb, e, t);
      while (true) {
        const bool last = e == ne;
        VT xv[kWinReps][4];
#pragma unroll
        for (int rep = 0; rep < kWinReps; ++rep)
#pragma unroll
          for (int u = 0; u < 4; ++u) xv[rep][u] = w.q[rep][u] >= 0 ? X[w.q[rep][u]] : (VT)0;
#pragma unroll
        for (int rep = 0; rep < kWinReps; ++rep) {
          AT* dst = prod + (rep * kWinBlock + t) * 4;
#pragma unroll
          for (int u = 0; u < 4; ++u) dst[u] = (AT)w.v[rep][u] * (AT)xv[rep][u];
        }
        __syncthreads();
        // The first pass's row data, loaded before the next window's entries
        // so that waiting for it does not wait for them. Row rc is always in
        // this window (it starts before e), so the first pass always runs.
        const int rf = rc + team;
        int rbf = 0, ref = 0;
        VT xf = (VT)0;
        double bf = 0.0;
        if (rf < r1) {
          rbf = indptr[rf];
          ref = indptr[rf + 1];
          if (MODE == 0 && partial) xf = X[rf];
          if (MODE == 1) bf = B[rf];
        }
        const int sn = e;
        const int en = last ? e : ((ne - e > kWin) ? e + kWin : ne);
        if (!last) win_load<MT, NTL>(w, indices, vals, sn, nb, en, t);
        // Rows that start before the window's end (in the last window, every
        // remaining row, empty trailing rows included). The valid rows form a
        // prefix of [rc, r1); exactly one of them, the last, reaches past e
        // or ends at e when this is not the last window.
        for (int base = rc;; base += kWinTeams) {
          const int r = base + team;
          int rb, re;
          VT xr;
          double br;
          if (base == rc) {
            rb = rbf;
            re = ref;
            xr = xf;
            br = bf;
          } else {
            if (base >= r1 || (!last && indptr[base] >= e)) break;
            rb = r < r1 ? indptr[r] : 0;
            re = r < r1 ? indptr[r + 1] : 0;
            xr = (MODE == 0 && partial && r < r1) ? X[r] : (VT)0;
            br = (MODE == 1 && r < r1) ? B[r] : 0.0;
          }
          const bool valid = r < r1 && (last || rb < e);
          AT acc = 0;
          if (valid) {
            const int lo = max(rb, s) - s;
            const int hi = min(re, e) - s;
            for (int i = lo + tl; i < hi; i += kRowLanes) acc += prod[i];
          }
#pragma unroll
          for (int o = kRowLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
          if (valid && tl == 0) {
            if (r == carry_row) acc = carry + acc;
            if (re > e) {
              s_carry = acc;
              s_next[0] = r;
              s_next[1] = r;
            } else {
              if (!last && re == e) {
                s_next[0] = r + 1;
                s_next[1] = -1;
              }
              if (MODE == 0) {
                const VT y = (VT)acc;
                Y[r] = y;
                if (partial) dot += (double)xr * (double)y;
              } else {
                const double y = br - (double)acc;
                Y[r] = (VT)y;
                dot += y * y;
              }
            }
          }
        }
        if (last) break;
        __syncthreads();
        rc = s_next[0];
        carry_row = s_next[1];
        carry = s_carry;
        s = sn;
        e = en;
        // The next window's products are written only after every thread has
        // read s_next and s_carry: they are rewritten only after its barrier.
      }
      // The next step's products must not overwrite this window's before
      // every team has summed its rows.
      __syncthreads();
    }
  }
  if (partial) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) dot += __shfl_xor(dot, o, 64);
    if (lane == 0) smem[wave] = dot;
    __syncthreads();
    if (t == 0) {
      double acc = 0.0;
      for (int v = 0; v < kWinBlock / 64; ++v) acc += smem[v];
      partial[b] = acc;
    }
  }
}

// x = 0, r = src * inv_scale[c], z = r/d, p = z; partials (r.z, r.r).
template <typename VT, int K>
__global__ __launch_bounds__(256) void k_cg_setup(int n, const double* __restrict__ src,
                                                  const double* __restrict__ inv_scale, VT* __restrict__ X,
                                                  VT* __restrict__ R, VT* __restrict__ Z, VT* __restrict__ P,
                                                  const double* __restrict__ diag, int precond,
                                                  double* __restrict__ partial) {
  __shared__ double smem[16 * 16];
  const size_t total = (size_t)n * K;
  const int c = threadIdx.x % K;
  const double s = inv_scale ? inv_scale[c] : 1.0;
  double rz = 0.0, rr = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const VT r = (VT)(src[e] * s);
    const double d = precond ? diag[e / K] : 0.0;
    const VT z = (precond && fabs(d) > 0.0) ? (VT)((double)r / d) : r;
    X[e] = (VT)0;
    R[e] = r;
    Z[e] = z;
    P[e] = z;
    rz += (double)r * (double)z;
    rr += (double)r * (double)r;
  }
  block_col_sum<K>(rz, partial + ((size_t)blockIdx.x * 2 + 0) * K, smem);
  block_col_sum<K>(rr, partial + ((size_t)blockIdx.x * 2 + 1) * K, smem);
}

// One block: ||b|| (0 -> 1), rho_0 = r.z, status. run_mask[c] == 0 marks a
// column that is not to be solved (already converged outside).
template <int K>
__global__ __launch_bounds__(256) void k_cg_setup_finish(const double* __restrict__ partial, int nparts,
                                                         ColState* __restrict__ st, const int* __restrict__ run_mask,
                                                         double* __restrict__ hist) {
  __shared__ double smem[256];
  __shared__ double sums[2 * K];
  reduce_partials<2 * K>(partial, nparts, sums, smem);
  if (threadIdx.x < K) {
    const int c = threadIdx.x;
    const double rz = sums[c];
    const double bb = sums[K + c];
    double bn = sqrt(bb);
    if (bn == 0.0) bn = 1.0;
    ColState s;
    s.rho = rz;
    s.bnorm = bn;
    s.rel = sqrt(bb) / bn;
    s.alpha = 0.0;
    s.beta = 0.0;
    s.status = (run_mask == nullptr || run_mask[c]) ? CC_RUNNING : CC_CONVERGED;
    s.iterations = 0;
    st[c] = s;
    if (hist) hist[c] = s.rel;
  }
}

// F1 (one block of kFinThreads).
template <int K>
__global__ __launch_bounds__(kFinThreads) void k_cg_alpha(const double* __restrict__ partial, int nparts,
                                                          ColState* __restrict__ st, int iter) {
  if (all_stopped<K>(st)) return;
  finalize_alpha<K>(partial, nparts, st, iter);
}

// U: x += alpha p, r -= alpha q, z = r/d; partials (r.r, r.z).
template <typename VT, int K>
__global__ __launch_bounds__(256) void k_cg_update(int n, VT* __restrict__ X, VT* __restrict__ R, VT* __restrict__ Z,
                                                   const VT* __restrict__ P, const VT* __restrict__ Q,
                                                   const double* __restrict__ diag, int precond,
                                                   const ColState* __restrict__ st,
                                                   double* __restrict__ partial_out) {
  __shared__ double smem[16 * 16];
  if (all_stopped<K>(st)) return;
  const int c = threadIdx.x % K;  // fixed per thread: the stride is a multiple of K
  const bool act = st[c].status == CC_RUNNING;
  const double a = st[c].alpha;
  const size_t total = (size_t)n * K;
  double rr = 0.0, rz = 0.0;
  if (act) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
      const VT x = (VT)((double)X[e] + a * (double)P[e]);
      const VT r = (VT)((double)R[e] - a * (double)Q[e]);
      const double d = precond ? diag[e / K] : 0.0;
      const VT z = (precond && fabs(d) > 0.0) ? (VT)((double)r / d) : r;
      X[e] = x;
      R[e] = r;
      Z[e] = z;
      rr += (double)r * (double)r;
      rz += (double)r * (double)z;
    }
  }
  block_col_sum<K>(rr, partial_out + ((size_t)blockIdx.x * 2 + 0) * K, smem);
  block_col_sum<K>(rz, partial_out + ((size_t)blockIdx.x * 2 + 1) * K, smem);
}

// F2 (one block of kFinThreads).
template <int K>
__global__ __launch_bounds__(kFinThreads) void k_cg_beta(const double* __restrict__ partial, int nparts,
                                                         ColState* __restrict__ st, int iter, double tol, int maxiter,
                                                         double* __restrict__ hist) {
  if (all_stopped<K>(st)) return;
  finalize_beta<K>(partial, nparts, st, iter, tol, maxiter, hist);
}

// P: p = z + beta p for the columns still running.
template <typename VT, int K>
__global__ __launch_bounds__(256) void k_cg_direction(int n, VT* __restrict__ P, const VT* __restrict__ Z,
                                                      const ColState* __restrict__ st) {
  if (all_stopped<K>(st)) return;
  const int c = threadIdx.x % K;
  if (st[c].status != CC_RUNNING) return;
  const double b = st[c].beta;
  const size_t total = (size_t)n * K;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
    P[e] = (VT)((double)Z[e] + b * (double)P[e]);
}

// Column sums of squares of a row-major [n][K] block: partial[block][c].
template <int K>
__global__ __launch_bounds__(256) void k_colsq(int n, const double* __restrict__ V, double* __restrict__ partial) {
  __shared__ double smem[16 * 16];
  const size_t total = (size_t)n * K;
  double s = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
    s += V[e] * V[e];
  block_col_sum<K>(s, partial + (size_t)blockIdx.x * K, smem);
}

// Outer refinement bookkeeping (one block): rel = ||B - A X|| / ||B||.
// first != 0: record ||B|| (0 -> 1). Sets per-column run flags and the scale
// of the next inner right-hand side (1 / ||residual||).
template <int K>
__global__ __launch_bounds__(256) void k_outer_finish(const double* __restrict__ partial, int nparts,
                                                      OuterState* __restrict__ ost, int* __restrict__ run_mask,
                                                      double* __restrict__ inv_scale, double tol, int first) {
  __shared__ double smem[256];
  __shared__ double sums[K];
  reduce_partials<K>(partial, nparts, sums, smem);
  if (threadIdx.x < K) {
    const int c = threadIdx.x;
    const double rn = sqrt(sums[c]);
    OuterState o = ost[c];
    if (first) {
      o.bnorm = rn == 0.0 ? 1.0 : rn;
      o.status = CC_RUNNING;
      o.outer = 0;
      o.inner_total = 0;
      o.best_rel = INFINITY;
    }
    o.rel = rn / o.bnorm;
    if (o.status == CC_RUNNING && o.rel <= tol) o.status = CC_CONVERGED;
    const bool run = o.status == CC_RUNNING;
    run_mask[c] = run;
    o.scale = rn;
    inv_scale[c] = (run && rn > 0.0) ? 1.0 / rn : 0.0;
    ost[c] = o;
  }
}

// X64 += scale[c] * D32 for the columns of this sweep.
template <int K>
__global__ __launch_bounds__(256) void k_outer_correct(int n, double* __restrict__ X, const float* __restrict__ D,
                                                       const OuterState* __restrict__ ost,
                                                       const int* __restrict__ run_mask) {
  const size_t total = (size_t)n * K;
  const int c = threadIdx.x % K;
  if (!run_mask[c]) return;
  const double s = ost[c].scale;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
    X[e] += s * (double)D[e];
}

template <int K>
__global__ void k_outer_account(OuterState* __restrict__ ost, const ColState* __restrict__ ist,
                                const int* __restrict__ run_mask) {
  const int c = threadIdx.x;
  if (c >= K || !run_mask[c]) return;
  ost[c].outer += 1;
  ost[c].inner_total += ist[c].iterations;
}

__global__ void k_zero(double* __restrict__ p, size_t count) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Host-side launch helpers.
// ---------------------------------------------------------------------------

static int stream_grid(size_t work) {
  size_t g = (work + kBlock * 4 - 1) / (kBlock * 4);
  if (g < 1) g = 1;
  if (g > (size_t)kMaxParts) g = kMaxParts;
  return (int)g;
}

template <typename VT, int K>
static int default_cpl() {
  const int w = 16 / (int)sizeof(VT);  // 16-byte vector loads
  return K >= w ? w : K;
}

template <int KL>
static int default_nz(double avg_row) {
  if (KL > 1) return 64 / KL;
  if (avg_row >= 96.0) return 64;
  if (avg_row >= 40.0) return 32;
  if (avg_row >= 20.0) return 16;
  return 8;
}

template <typename MT, typename VT, int K, int NZ, int MODE, int CPL>
static int spmm_launch_nz(int n, int unroll, int grid_req, int stream, int nsub, const int* indptr,
                          const int* indices, const MT* vals, const VT* X, VT* Y, const double* B, double* partial,
                          const ColState* st, hipStream_t s, int* grid_out) {
  constexpr int T = (K / CPL) * NZ;
  const int teams_per_block = (kBlock / 64) * (64 / T);
  int grid = grid_req;
  if (grid <= 0) grid = (n + teams_per_block - 1) / teams_per_block;
  if (grid > kMaxParts) grid = kMaxParts;
  if (grid >= 2 * kXcds) grid -= grid % kXcds;
  if (grid < 1) grid = 1;
  if (unroll == 8 && stream)
    launch_timed(k_spmm<MT, VT, K, NZ, MODE, 8, CPL, true>, dim3(grid), dim3(kBlock), 0, s, n, indptr, indices,
                 vals, X, Y, B, partial, st, nsub);
  else if (unroll == 8)
    launch_timed(k_spmm<MT, VT, K, NZ, MODE, 8, CPL, false>, dim3(grid), dim3(kBlock), 0, s, n, indptr,
                 indices, vals, X, Y, B, partial, st, nsub);
  else if (unroll >= 4 && stream)
    launch_timed(k_spmm<MT, VT, K, NZ, MODE, 4, CPL, true>, dim3(grid), dim3(kBlock), 0, s, n, indptr, indices,
                 vals, X, Y, B, partial, st, nsub);
  else if (unroll >= 4)
    launch_timed(k_spmm<MT, VT, K, NZ, MODE, 4, CPL, false>, dim3(grid), dim3(kBlock), 0, s, n, indptr,
                 indices, vals, X, Y, B, partial, st, nsub);
  else
    launch_timed(k_spmm<MT, VT, K, NZ, MODE, 1, CPL, false>, dim3(grid), dim3(kBlock), 0, s, n, indptr,
                 indices, vals, X, Y, B, partial, st, nsub);
  if (grid_out) *grid_out = grid;
  return (int)hipGetLastError();
}

template <typename MT, typename VT, int K, int MODE, int CPL>
static int spmm_launch_cpl(int n, double avg_row, const int* indptr, const int* indices, const MT* vals, const VT* X,
                           VT* Y, const double* B, double* partial, const ColState* st, hipStream_t s, int* grid_out) {
  constexpr int KL = K / CPL;
  const SpmmConfig cfg = spmm_cfg<VT>(K);
  // A shape tuned for another vector width may not fit this one's teams.
  const int nz = (cfg.nz > 0 && KL * cfg.nz <= 64) ? cfg.nz : default_nz<KL>(avg_row);
  const int unroll = cfg.unroll > 0 ? cfg.unroll : 4;
  const int grid = cfg.grid;
  const int nsub = cfg.nsub > 0 ? cfg.nsub : 1;
#define CC_NZ(NZV)                                                                                                  \
  if (nz == NZV && KL * NZV <= 64)                                                                                  \
    return spmm_launch_nz<MT, VT, K, (KL * NZV <= 64 ? NZV : 1), MODE, CPL>(n, unroll, grid, cfg.stream, nsub,      \
                                                                          indptr, indices, vals, X, Y, B, partial, \
                                                                          st, s, grid_out);
  CC_NZ(64)
  CC_NZ(32)
  CC_NZ(16)
  CC_NZ(8)
  CC_NZ(4)
  CC_NZ(2)
  CC_NZ(1)
#undef CC_NZ
  return CC_E_ARG;
}

template <typename MT, typename VT, int MODE>
static int spmv_win_launch(int n, double avg_row, int grid_req, int stream, int nsub, const int* indptr,
                           const int* indices, const MT* vals, const VT* X, VT* Y, const double* B, double* partial,
                           const ColState* st, hipStream_t s, int* grid_out) {
  int grid = grid_req;
  if (grid <= 0) grid = (int)(avg_row * n / (2.0 * kWin)) + 1;
  if (grid > kMaxParts) grid = kMaxParts;
  if (grid >= 2 * kXcds) grid -= grid % kXcds;
  if (grid < 1) grid = 1;
  if (stream)
    launch_timed(k_spmv_win<MT, VT, MODE, true>, dim3(grid), dim3(kWinBlock), 0, s, n, indptr, indices, vals,
                 X, Y, B, partial, st, nsub);
  else
    launch_timed(k_spmv_win<MT, VT, MODE, false>, dim3(grid), dim3(kWinBlock), 0, s, n, indptr, indices, vals,
                 X, Y, B, partial, st, nsub);
  if (grid_out) *grid_out = grid;
  return (int)hipGetLastError();
}

template <typename MT, typename VT, int K, int MODE>
static int spmm_launch(int n, double avg_row, const int* indptr, const int* indices, const MT* vals, const VT* X,
                       VT* Y, const double* B, double* partial, const ColState* st, hipStream_t s, int* grid_out) {
  if constexpr (K == 1) {
    // The window kernel loads 4 entries at a time with 16-byte loads.
    const SpmmConfig& cfg = spmm_cfg<VT>(1);
    if (cfg.unroll == kUnrollWindow && ((uintptr_t)indices & 15) == 0 && ((uintptr_t)vals & 15) == 0)
      return spmv_win_launch<MT, VT, MODE>(n, avg_row, cfg.grid, cfg.stream, cfg.nsub > 0 ? cfg.nsub : 1, indptr,
                                           indices, vals, X, Y, B, partial, st, s, grid_out);
  }
  int cpl = spmm_cfg<VT>(K).cpl > 0 ? spmm_cfg<VT>(K).cpl : default_cpl<VT, K>();
  if (cpl > K || cpl * (int)sizeof(VT) > 16) cpl = default_cpl<VT, K>();
  if (cpl >= 4 && sizeof(VT) == 4 && K >= 4)
    return spmm_launch_cpl<MT, VT, K, MODE, (sizeof(VT) == 4 && K >= 4 ? 4 : 1)>(n, avg_row, indptr, indices, vals, X,
                                                                                 Y, B, partial, st, s, grid_out);
  if (cpl >= 2 && K >= 2)
    return spmm_launch_cpl<MT, VT, K, MODE, (K >= 2 ? 2 : 1)>(n, avg_row, indptr, indices, vals, X, Y, B, partial, st,
                                                              s, grid_out);
  return spmm_launch_cpl<MT, VT, K, MODE, 1>(n, avg_row, indptr, indices, vals, X, Y, B, partial, st, s, grid_out);
}

static double avg_row_len(int n, const int* indptr, hipStream_t s) {
  int ends[2] = {0, 0};
  if (hipMemcpyAsync(&ends[0], indptr, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return 64.0;
  if (hipMemcpyAsync(&ends[1], indptr + n, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) return 64.0;
  if (hipStreamSynchronize(s) != hipSuccess) return 64.0;
  return n > 0 ? double(ends[1] - ends[0]) / n : 0.0;
}

// The mean row length only chooses a default launch shape; reading it costs
// a stream synchronisation, so the single-product entry points skip it when
// the configured shape (set by the autotuner) does not need it.
template <typename VT>
static double avg_row_if_needed(int n, int k, const int* indptr, hipStream_t s) {
  const SpmmConfig& cfg = spmm_cfg<VT>(k);
  if (cfg.grid > 0 && (cfg.nz > 0 || (k == 1 && cfg.unroll == kUnrollWindow))) return 0.0;
  return avg_row_len(n, indptr, s);
}

// Workspace carving.
struct Carve {
  char* base;
  size_t off = 0;
  size_t cap;
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = reinterpret_cast<T*>(base + off);
    off += count * sizeof(T);
    return p;
  }
  bool ok() const { return off <= cap; }
};

template <typename VT>
struct CgBuffers {
  VT *X, *R, *Z, *P, *Q;
  double *diag, *pa, *pb;
  ColState* st;
};

// tail: extra rows after P and Q (an operator may keep per-apply data
// there, as the E-fused product keeps its small dofs' values).
template <typename VT>
static size_t cg_bytes(size_t n, int k, size_t tail = 0) {
  size_t b = 0;
  auto add = [&](size_t bytes) { b = ((b + 255) & ~size_t(255)) + bytes; };
  for (int i = 0; i < 5; ++i) add((i >= 3 ? n + tail : n) * k * sizeof(VT));
  add(n * sizeof(double));
  add((size_t)kMaxParts * 2 * k * sizeof(double));
  add((size_t)kMaxParts * 2 * k * sizeof(double));
  add(k * sizeof(ColState));
  return b + 256;
}

template <typename VT>
static CgBuffers<VT> carve_cg(Carve& cv, size_t n, int k, VT* X_external, size_t tail = 0) {
  CgBuffers<VT> b;
  b.X = cv.take<VT>(n * k);
  if (X_external) b.X = X_external;
  b.R = cv.take<VT>(n * k);
  b.Z = cv.take<VT>(n * k);
  b.P = cv.take<VT>((n + tail) * k);
  b.Q = cv.take<VT>((n + tail) * k);
  b.diag = cv.take<double>(n);
  b.pa = cv.take<double>((size_t)kMaxParts * 2 * k);
  b.pb = cv.take<double>((size_t)kMaxParts * 2 * k);
  b.st = cv.take<ColState>(k);
  return b;
}

// The device-resident CG loop for K columns over any operator:
// apply(P, Q, partial, &nparts) launches Q = A P and writes the per-block
// partials of P.Q. src is the float64 right-hand side block; inv_scale (device, may be
// null) scales it into VT. The per-column state is left in b.st (and copied
// to host_st).
// after_iter (may be null) is called with the iteration number once that
// iteration's launches are enqueued (it may enqueue copies of b.X).
template <typename VT, int K, typename ApplyFn>
static int cg_loop(int n, ApplyFn&& apply, const double* src, const double* inv_scale, const int* run_mask,
                   CgBuffers<VT>& b, int precond, double tol, int maxiter, int check_every, hipStream_t s,
                   std::vector<ColState>& host_st, double* hist,
                   const std::function<int(int)>* after_iter = nullptr, bool one_launch = false) {
  const int sg = stream_grid((size_t)n * K);
  hipLaunchKernelGGL((k_cg_setup<VT, K>), dim3(sg), dim3(kBlock), 0, s, n, src, inv_scale, b.X, b.R, b.Z, b.P, b.diag,
                     precond, b.pa);
  hipLaunchKernelGGL((k_cg_setup_finish<K>), dim3(1), dim3(kBlock), 0, s, b.pa, sg, b.st, run_mask, hist);
  CC_CHECK(hipGetLastError());
  if (check_every < 1) check_every = 1;
  host_st.resize(K);
  const size_t pmark = prof_pending_mark();
  for (int it = 1; it <= maxiter; ++it) {
    int gs = 0;
    // A product that is one launch is timed by that launch's own events.
    if (one_launch) {
      prof_arm(kProfSpmm);
      CC_CHECK(apply(b.P, b.Q, b.pa, &gs));
      prof_disarm();
    } else {
      prof_begin(kProfSpmm, s);
      CC_CHECK(apply(b.P, b.Q, b.pa, &gs));
      prof_end(s);
    }
    hipLaunchKernelGGL((k_cg_alpha<K>), dim3(1), dim3(kFinThreads), 0, s, b.pa, gs, b.st, it);
    prof_begin(kProfUpdate, s);
    hipLaunchKernelGGL((k_cg_update<VT, K>), dim3(sg), dim3(kBlock), 0, s, n, b.X, b.R, b.Z, b.P, b.Q, b.diag,
                       precond, b.st, b.pb);
    prof_end(s);
    hipLaunchKernelGGL((k_cg_beta<K>), dim3(1), dim3(kFinThreads), 0, s, b.pb, sg, b.st, it, tol, maxiter, hist);
    prof_begin(kProfDirection, s);
    hipLaunchKernelGGL((k_cg_direction<VT, K>), dim3(sg), dim3(kBlock), 0, s, n, b.P, b.Z, b.st);
    prof_end(s);
    CC_CHECK(hipGetLastError());
    if (after_iter) CC_CHECK((*after_iter)(it));
    if (it % check_every == 0 || it == maxiter) {
      CC_CHECK(hipMemcpyAsync(host_st.data(), b.st, K * sizeof(ColState), hipMemcpyDeviceToHost, s));
      CC_CHECK(hipStreamSynchronize(s));
      bool any = false;
      for (int c = 0; c < K; ++c) any |= host_st[c].status == CC_RUNNING;
      if (!any) break;
    }
  }
  CC_CHECK(hipMemcpyAsync(host_st.data(), b.st, K * sizeof(ColState), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  // Iteration it's product ran while some column was running: it <= the
  // last column's count.
  int live = 0;
  for (int c = 0; c < K; ++c) live = std::max(live, (int)host_st[c].iterations);
  prof_keep(kProfSpmm, pmark, live);
  prof_collect();
  return 0;
}

// The runs of the rows (runs.hip), for single-right-hand-side products.
struct RunsView {
  const int* rptr;
  const uint32_t* runs;
};
template <typename MT, typename VT>
int runs_spmv_cg(int n, const int* rptr, const uint32_t* runs, const int* indptr, const MT* vals, const VT* P, VT* Q,
                 double* partial, const ColState* st, hipStream_t s, int* grid_out);

// rb (may be NULL): take the products from the row-block form instead;
// rv (may be NULL, K = 1 only): from the run encoding of the rows.
template <typename MT, typename VT, int K>
static int cg_run(int n, double avg_row, const int* indptr, const int* indices, const MT* vals, const double* src,
                  const double* inv_scale, const int* run_mask, CgBuffers<VT>& b, int precond, double tol,
                  int maxiter, int check_every, hipStream_t s, std::vector<ColState>& host_st,
                  double* hist = nullptr, const cc_rb_desc* rb = nullptr, const RunsView* rv = nullptr) {
  const ColState* st = b.st;
  auto apply = [&](const VT* P, VT* Q, double* partial, int* gs) {
    if constexpr (K == 1) {
      if (rv) return runs_spmv_cg<MT, VT>(n, rv->rptr, rv->runs, indptr, vals, P, Q, partial, st, s, gs);
    }
    if (rb) return rb_spmm<MT, VT, K>(*rb, P, Q, partial, st, s, gs);
    return spmm_launch<MT, VT, K, 0>(n, avg_row, indptr, indices, vals, P, Q, nullptr, partial, st, s, gs);
  };
  return cg_loop<VT, K>(n, apply, src, inv_scale, run_mask, b, precond, tol, maxiter, check_every, s, host_st, hist,
                        nullptr, true);
}

// Generic float64 single-right-hand-side CG over an operator (used by the
// matrix-free operator in sumfact.hip). diag_src (device) is copied in as the
// Jacobi diagonal. The float64 true residual is computed by `residual`.
int cg_operator_f64(int n, const std::function<int(const double*, double*, double*, int*, const ColState*)>& apply,
                    const double* diag_src, const double* B, double* X, double tol, int maxiter, int precond,
                    int check_every, void* ws, size_t wsb, cc_pcg_result* result, double* hist, hipStream_t s,
                    const std::function<int(const double*, double*, double*)>& residual, const int* snap_iters,
                    int n_snap, double* snapshots) {
  Carve cv{(char*)ws, 0, wsb};
  CgBuffers<double> b = carve_cg<double>(cv, n, 1, X);
  if (!cv.ok()) return CC_E_WORKSPACE;
  if (precond && diag_src)
    CC_CHECK(hipMemcpyAsync(b.diag, diag_src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
  const ColState* st = b.st;
  auto ap = [&](const double* P, double* Q, double* partial, int* gs) { return apply(P, Q, partial, gs, st); };
  std::vector<ColState> hs;
  // Snapshot j is the iterate after snap_iters[j] iterations (or the final
  // iterate when the solve stopped earlier).
  std::function<int(int)> snap = [&](int it) {
    for (int j = 0; j < n_snap; ++j)
      if (snap_iters[j] == it)
        CC_CHECK(hipMemcpyAsync(snapshots + (size_t)j * n, b.X, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice,
                                s));
    return 0;
  };
  CC_CHECK((cg_loop<double, 1>(n, ap, B, nullptr, nullptr, b, precond && diag_src, tol, maxiter, check_every, s, hs,
                               hist, n_snap > 0 ? &snap : nullptr)));
  for (int j = 0; j < n_snap; ++j)
    if (snap_iters[j] > hs[0].iterations)
      CC_CHECK(hipMemcpyAsync(snapshots + (size_t)j * n, b.X, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
  CC_CHECK(hipStreamSynchronize(s));
  double rel = 0.0;
  CC_CHECK(residual(X, b.R, &rel));
  result->iterations = hs[0].iterations;
  result->status = hs[0].status == CC_RUNNING ? CC_MAXITER : hs[0].status;
  result->outer = 1;
  result->inner_total = hs[0].iterations;
  result->recursive_rel = hs[0].rel;
  result->true_rel = rel;
  return 0;
}

size_t cg_operator_bytes(size_t n) { return cg_bytes<double>(n, 1); }

// Float64 Jacobi-PCG for K right-hand sides sharing an operator (row-major
// [n][K] blocks, as the CSR solver's): apply(P, Q, partial, &nparts) writes
// Q = A P and per-block, per-column partials of P.Q ([nparts][K]);
// residual(X, scratch, rel) writes the K true relative residuals (host).
template <int K>
static int cg_operator_k(int n, const std::function<int(const double*, double*, double*, int*)>& apply,
                         const double* diag_src, const double* B, double* X, double tol, int maxiter, int precond,
                         int check_every, void* ws, size_t wsb, cc_pcg_result* results, hipStream_t s,
                         const std::function<int(const double*, double*, double*)>& residual, size_t tail) {
  Carve cv{(char*)ws, 0, wsb};
  CgBuffers<double> b = carve_cg<double>(cv, n, K, X, tail);
  if (!cv.ok()) return CC_E_WORKSPACE;
  if (precond && diag_src)
    CC_CHECK(hipMemcpyAsync(b.diag, diag_src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
  std::vector<ColState> hs;
  CC_CHECK((cg_loop<double, K>(n, apply, B, nullptr, nullptr, b, precond && diag_src, tol, maxiter, check_every, s,
                               hs, nullptr)));
  CC_CHECK(hipStreamSynchronize(s));
  std::vector<double> rel(K);
  CC_CHECK(residual(X, b.R, rel.data()));
  for (int c = 0; c < K; ++c) {
    results[c].iterations = hs[c].iterations;
    results[c].status = hs[c].status == CC_RUNNING ? CC_MAXITER : hs[c].status;
    results[c].outer = 1;
    results[c].inner_total = hs[c].iterations;
    results[c].recursive_rel = hs[c].rel;
    results[c].true_rel = rel[c];
  }
  return 0;
}

int cg_operator_multi(int n, int k, const std::function<int(const double*, double*, double*, int*)>& apply,
                      const double* diag_src, const double* B, double* X, double tol, int maxiter, int precond,
                      int check_every, void* ws, size_t wsb, cc_pcg_result* results, hipStream_t s,
                      const std::function<int(const double*, double*, double*)>& residual, size_t tail) {
#define CC_OPK(KK)                                                                                            \
  case KK:                                                                                                    \
    return cg_operator_k<KK>(n, apply, diag_src, B, X, tol, maxiter, precond, check_every, ws, wsb, results, s, \
                             residual, tail);
  switch (k) {
    CC_OPK(1)
    CC_OPK(2)
    CC_OPK(4)
    CC_OPK(8)
    CC_OPK(16)
    default: return CC_E_ARG;
  }
#undef CC_OPK
}

size_t cg_operator_multi_bytes(size_t n, int k, size_t tail) { return cg_bytes<double>(n, k, tail); }

// Y = A X over n rows for k columns with the autotuned CSR launch shape
// (float64); partial[block][c] receives the per-block sums of X.Y when not
// null, *gs the number of blocks. avg_row (the mean row length) picks the
// default shape when the autotuner set none.
int csr_spmm_f64(int n, int k, double avg_row, const int* indptr, const int* indices, const double* vals,
                 const double* X, double* Y, double* partial, hipStream_t s, int* gs) {
#define CC_SPK(KK)                                                                                                 \
  case KK:                                                                                                         \
    return spmm_launch<double, double, KK, 0>(n, avg_row, indptr, indices, vals, X, Y, nullptr, partial, nullptr, s, \
                                              gs);
  switch (k) {
    CC_SPK(1)
    CC_SPK(2)
    CC_SPK(4)
    CC_SPK(8)
    CC_SPK(16)
    default: return CC_E_ARG;
  }
#undef CC_SPK
}

// True relative residuals ||B - A X|| / ||B|| (||B|| = 0 -> 1) of the float64
// block X, written to rel[K] (host). The per-block partial sums are combined
// on the host in block order.
template <int K>
static int true_residual(int n, double avg_row, const int* indptr, const int* indices, const double* vals,
                         const double* B, const double* X, double* R, double* partial, hipStream_t s, double* rel) {
  int gs = 0;
  CC_CHECK((spmm_launch<double, double, K, 1>(n, avg_row, indptr, indices, vals, X, R, B, partial, nullptr, s, &gs)));
  const int sg = stream_grid((size_t)n * K);
  hipLaunchKernelGGL((k_colsq<K>), dim3(sg), dim3(kBlock), 0, s, n, B, partial + (size_t)kMaxParts * K);
  CC_CHECK(hipGetLastError());
  std::vector<double> hr((size_t)gs * K), hb((size_t)sg * K);
  CC_CHECK(hipMemcpyAsync(hr.data(), partial, hr.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipMemcpyAsync(hb.data(), partial + (size_t)kMaxParts * K, hb.size() * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  for (int c = 0; c < K; ++c) {
    double rr = 0.0, bb = 0.0;
    for (int g = 0; g < gs; ++g) rr += hr[(size_t)g * K + c];
    for (int g = 0; g < sg; ++g) bb += hb[(size_t)g * K + c];
    double bn = std::sqrt(bb);
    if (bn == 0.0) bn = 1.0;
    rel[c] = std::sqrt(rr) / bn;
  }
  return 0;
}

template <int K>
static int pcg_f64(int n, const int* indptr, const int* indices, const double* vals, const double* B, double* X,
                   double tol, int maxiter, int precond, int check_every, void* ws, size_t wsb,
                   cc_pcg_result* results, double* hist, hipStream_t s, const cc_rb_desc* rb = nullptr,
                   const RunsView* rv = nullptr) {
  Carve cv{(char*)ws, 0, wsb};
  CgBuffers<double> b = carve_cg<double>(cv, n, K, X);
  if (!cv.ok()) return CC_E_WORKSPACE;
  const double avg = avg_row_len(n, indptr, s);
  if (precond) {
    hipLaunchKernelGGL(k_diag, dim3((n + 255) / 256), dim3(256), 0, s, n, indptr, indices, vals, b.diag);
    CC_CHECK(hipGetLastError());
  }
  std::vector<ColState> hs;
  CC_CHECK((cg_run<double, double, K>(n, avg, indptr, indices, vals, B, nullptr, nullptr, b, precond, tol, maxiter,
                                      check_every, s, hs, hist, rb, rv)));
  std::vector<double> rel(K);
  CC_CHECK(true_residual<K>(n, avg, indptr, indices, vals, B, X, b.R, b.pa, s, rel.data()));
  for (int c = 0; c < K; ++c) {
    results[c].iterations = hs[c].iterations;
    results[c].status = hs[c].status == CC_RUNNING ? CC_MAXITER : hs[c].status;
    results[c].outer = 1;
    results[c].inner_total = hs[c].iterations;
    results[c].recursive_rel = hs[c].rel;
    results[c].true_rel = rel[c];
  }
  return 0;
}

template <int K>
static size_t mixed_bytes(size_t n) {
  size_t b = cg_bytes<float>(n, K);
  auto add = [&](size_t bytes) { b = ((b + 255) & ~size_t(255)) + bytes; };
  add(n * K * sizeof(double));                  // R64
  add((size_t)kMaxParts * K * sizeof(double));  // residual partials
  add(K * sizeof(OuterState));
  add(K * sizeof(int));
  add(K * sizeof(double));
  return b + 256;
}

template <int K>
static int pcg_mixed(int n, const int* indptr, const int* indices, const double* vals, const float* vals32,
                     const double* B, double* X, double tol, int maxiter, double inner_tol, int max_outer,
                     int precond, int check_every, void* ws, size_t wsb, cc_pcg_result* results, hipStream_t s,
                     const cc_rb_desc* rb = nullptr, const RunsView* rv = nullptr) {
  Carve cv{(char*)ws, 0, wsb};
  CgBuffers<float> b = carve_cg<float>(cv, n, K, nullptr);
  double* R64 = cv.take<double>((size_t)n * K);
  double* pr = cv.take<double>((size_t)kMaxParts * K);
  OuterState* ost = cv.take<OuterState>(K);
  int* run_mask = cv.take<int>(K);
  double* inv_scale = cv.take<double>(K);
  if (!cv.ok()) return CC_E_WORKSPACE;
  const double avg = avg_row_len(n, indptr, s);
  if (precond) {
    hipLaunchKernelGGL(k_diag, dim3((n + 255) / 256), dim3(256), 0, s, n, indptr, indices, vals, b.diag);
  }
  const int sg = stream_grid((size_t)n * K);
  hipLaunchKernelGGL(k_zero, dim3(sg), dim3(kBlock), 0, s, X, (size_t)n * K);
  int gs = 0;
  // R64 = B - A X in float64 with partials of R64.R64: from the float64
  // row-block form when there is one (0.6 of the CSR product's time on the
  // 3D systems), else from the CSR arrays.
  auto residual64 = [&]() -> int {
    if (rb && rb->gval && !rb->pairs) return rb_spmm<double, double, K>(*rb, X, R64, pr, nullptr, s, &gs, B);
    return spmm_launch<double, double, K, 1>(n, avg, indptr, indices, vals, X, R64, B, pr, nullptr, s, &gs);
  };
  CC_CHECK(residual64());
  hipLaunchKernelGGL((k_outer_finish<K>), dim3(1), dim3(kBlock), 0, s, pr, gs, ost, run_mask, inv_scale, tol, 1);
  CC_CHECK(hipGetLastError());
  std::vector<OuterState> ho(K);
  std::vector<ColState> hs;
  int used = 0;
  for (int sweep = 0; sweep < max_outer; ++sweep) {
    CC_CHECK(hipMemcpyAsync(ho.data(), ost, K * sizeof(OuterState), hipMemcpyDeviceToHost, s));
    CC_CHECK(hipStreamSynchronize(s));
    bool any = false;
    for (int c = 0; c < K; ++c) any |= ho[c].status == CC_RUNNING;
    if (!any || used >= maxiter) break;
    CC_CHECK((cg_run<float, float, K>(n, avg, indptr, indices, vals32, R64, inv_scale, run_mask, b, precond,
                                      inner_tol, maxiter - used, check_every, s, hs, nullptr, rb, rv)));
    int most = 0;
    for (int c = 0; c < K; ++c)
      if (ho[c].status == CC_RUNNING && hs[c].iterations > most) most = hs[c].iterations;
    used += most > 0 ? most : 1;
    hipLaunchKernelGGL((k_outer_correct<K>), dim3(sg), dim3(kBlock), 0, s, n, X, b.X, ost, run_mask);
    hipLaunchKernelGGL((k_outer_account<K>), dim3(1), dim3(64), 0, s, ost, b.st, run_mask);
    CC_CHECK(residual64());
    hipLaunchKernelGGL((k_outer_finish<K>), dim3(1), dim3(kBlock), 0, s, pr, gs, ost, run_mask, inv_scale, tol, 0);
    CC_CHECK(hipGetLastError());
  }
  CC_CHECK(hipMemcpyAsync(ho.data(), ost, K * sizeof(OuterState), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  for (int c = 0; c < K; ++c) {
    results[c].iterations = ho[c].inner_total;
    results[c].status = ho[c].status == CC_CONVERGED ? CC_CONVERGED : CC_MAXITER;
    results[c].outer = ho[c].outer;
    results[c].inner_total = ho[c].inner_total;
    results[c].recursive_rel = ho[c].rel;
    results[c].true_rel = ho[c].rel;
  }
  return 0;
}

// FP32 inner / FP64 outer refinement over an operator, as pcg_mixed over a
// CSR matrix: residual(X, R64, partial, &nparts) writes R64 = B - A X in
// float64 with per-block, per-column partials of R64.R64 ([nparts][K],
// nparts <= kMaxParts); apply32(P, Q, partial, &nparts) is the float32
// operator with per-block partials of P.Q. The inner solves start from zero
// on the scaled residual and stop at inner_tol; the outer loop stops on the
// float64 residual.
template <int K>
static int mixed_operator_k(int n, const std::function<int(const double*, double*, double*, int*)>& residual,
                            const std::function<int(const float*, float*, double*, int*)>& apply32,
                            const double* diag_src, double* X, double tol, int maxiter, double inner_tol,
                            int max_outer, int precond, int check_every, void* ws, size_t wsb,
                            cc_pcg_result* results, hipStream_t s) {
  Carve cv{(char*)ws, 0, wsb};
  CgBuffers<float> b = carve_cg<float>(cv, n, K, nullptr);
  double* R64 = cv.take<double>((size_t)n * K);
  double* pr = cv.take<double>((size_t)kMaxParts * K);
  OuterState* ost = cv.take<OuterState>(K);
  int* run_mask = cv.take<int>(K);
  double* inv_scale = cv.take<double>(K);
  if (!cv.ok()) return CC_E_WORKSPACE;
  const int pc = precond && diag_src ? 1 : 0;
  if (pc) CC_CHECK(hipMemcpyAsync(b.diag, diag_src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
  const int sg = stream_grid((size_t)n * K);
  hipLaunchKernelGGL(k_zero, dim3(sg), dim3(kBlock), 0, s, X, (size_t)n * K);
  int gs = 0;
  CC_CHECK(residual(X, R64, pr, &gs));
  if (gs < 1 || gs > kMaxParts) return CC_E_ARG;
  hipLaunchKernelGGL((k_outer_finish<K>), dim3(1), dim3(kBlock), 0, s, pr, gs, ost, run_mask, inv_scale, tol, 1);
  CC_CHECK(hipGetLastError());
  std::vector<OuterState> ho(K);
  std::vector<ColState> hs;
  int used = 0;
  for (int sweep = 0; sweep < max_outer; ++sweep) {
    CC_CHECK(hipMemcpyAsync(ho.data(), ost, K * sizeof(OuterState), hipMemcpyDeviceToHost, s));
    CC_CHECK(hipStreamSynchronize(s));
    bool any = false;
    for (int c = 0; c < K; ++c) any |= ho[c].status == CC_RUNNING;
    if (!any || used >= maxiter) break;
    CC_CHECK((cg_loop<float, K>(n, apply32, R64, inv_scale, run_mask, b, pc, inner_tol, maxiter - used, check_every,
                                s, hs, nullptr)));
    int most = 0;
    for (int c = 0; c < K; ++c)
      if (ho[c].status == CC_RUNNING && hs[c].iterations > most) most = hs[c].iterations;
    used += most > 0 ? most : 1;
    hipLaunchKernelGGL((k_outer_correct<K>), dim3(sg), dim3(kBlock), 0, s, n, X, b.X, ost, run_mask);
    hipLaunchKernelGGL((k_outer_account<K>), dim3(1), dim3(64), 0, s, ost, b.st, run_mask);
    CC_CHECK(residual(X, R64, pr, &gs));
    if (gs < 1 || gs > kMaxParts) return CC_E_ARG;
    hipLaunchKernelGGL((k_outer_finish<K>), dim3(1), dim3(kBlock), 0, s, pr, gs, ost, run_mask, inv_scale, tol, 0);
    CC_CHECK(hipGetLastError());
  }
  CC_CHECK(hipMemcpyAsync(ho.data(), ost, K * sizeof(OuterState), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  for (int c = 0; c < K; ++c) {
    results[c].iterations = ho[c].inner_total;
    results[c].status = ho[c].status == CC_CONVERGED ? CC_CONVERGED : CC_MAXITER;
    results[c].outer = ho[c].outer;
    results[c].inner_total = ho[c].inner_total;
    results[c].recursive_rel = ho[c].rel;
    results[c].true_rel = ho[c].rel;
  }
  return 0;
}

// Bs = R * inv_scale[c] (zero for the columns not running: inv_scale 0).
template <int K>
__global__ __launch_bounds__(256) void k_outer_scale(int n, const double* __restrict__ R,
                                                     const double* __restrict__ inv_scale, double* __restrict__ Bs) {
  const size_t total = (size_t)n * K;
  const double sc = inv_scale[threadIdx.x % K];
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
    Bs[e] = R[e] * sc;
}

// X += scale[c] * Z for the columns of this sweep (float64 correction).
template <int K>
__global__ __launch_bounds__(256) void k_outer_correct64(int n, double* __restrict__ X, const double* __restrict__ Z,
                                                         const OuterState* __restrict__ ost,
                                                         const int* __restrict__ run_mask) {
  const size_t total = (size_t)n * K;
  const int c = threadIdx.x % K;
  if (!run_mask[c]) return;
  const double sc = ost[c].scale;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
    X[e] += sc * Z[e];
}

template <int K>
static size_t refine_bytes(size_t n) {
  size_t b = 0;
  auto add = [&](size_t bytes) { b = ((b + 255) & ~size_t(255)) + bytes; };
  for (int i = 0; i < 3; ++i) add(n * K * sizeof(double));  // R64, Bs, Z
  add((size_t)kMaxParts * K * sizeof(double));
  add(K * sizeof(OuterState));
  add(K * sizeof(int));
  add(K * sizeof(double));
  return b + 256;
}

// Iterative refinement around an inner solver: the outer loop as in
// mixed_operator_k (float64 residual, per-column stop and scale), the inner
// solve inner(Bs, Z, maxiter_left, its) on the scaled residual Bs, with its
// per-column iteration counts written to its (host).
template <int K>
static int refine_operator_k(int n, const std::function<int(const double*, double*, double*, int*)>& residual,
                             const std::function<int(const double*, double*, int, int*)>& inner, double* X,
                             double tol, int maxiter, int max_outer, void* ws, size_t wsb, cc_pcg_result* results,
                             hipStream_t s) {
  Carve cv{(char*)ws, 0, wsb};
  double* R64 = cv.take<double>((size_t)n * K);
  double* Bs = cv.take<double>((size_t)n * K);
  double* Z = cv.take<double>((size_t)n * K);
  double* pr = cv.take<double>((size_t)kMaxParts * K);
  OuterState* ost = cv.take<OuterState>(K);
  int* run_mask = cv.take<int>(K);
  double* inv_scale = cv.take<double>(K);
  if (!cv.ok()) return CC_E_WORKSPACE;
  const int sg = stream_grid((size_t)n * K);
  hipLaunchKernelGGL(k_zero, dim3(sg), dim3(kBlock), 0, s, X, (size_t)n * K);
  int gs = 0;
  CC_CHECK(residual(X, R64, pr, &gs));
  if (gs < 1 || gs > kMaxParts) return CC_E_ARG;
  hipLaunchKernelGGL((k_outer_finish<K>), dim3(1), dim3(kBlock), 0, s, pr, gs, ost, run_mask, inv_scale, tol, 1);
  CC_CHECK(hipGetLastError());
  std::vector<OuterState> ho(K);
  std::vector<int> outer(K, 0), inner_total(K, 0), its(K, 0);
  int used = 0;
  for (int sweep = 0; sweep < max_outer; ++sweep) {
    CC_CHECK(hipMemcpyAsync(ho.data(), ost, K * sizeof(OuterState), hipMemcpyDeviceToHost, s));
    CC_CHECK(hipStreamSynchronize(s));
    bool any = false;
    for (int c = 0; c < K; ++c) any |= ho[c].status == CC_RUNNING;
    if (!any || used >= maxiter) break;
    hipLaunchKernelGGL((k_outer_scale<K>), dim3(sg), dim3(kBlock), 0, s, n, R64, inv_scale, Bs);
    CC_CHECK(hipGetLastError());
    std::fill(its.begin(), its.end(), 0);
    CC_CHECK(inner(Bs, Z, maxiter - used, its.data()));
    int most = 0;
    for (int c = 0; c < K; ++c)
      if (ho[c].status == CC_RUNNING) {
        outer[c] += 1;
        inner_total[c] += its[c];
        most = its[c] > most ? its[c] : most;
      }
    used += most > 0 ? most : 1;
    hipLaunchKernelGGL((k_outer_correct64<K>), dim3(sg), dim3(kBlock), 0, s, n, X, Z, ost, run_mask);
    CC_CHECK(residual(X, R64, pr, &gs));
    if (gs < 1 || gs > kMaxParts) return CC_E_ARG;
    hipLaunchKernelGGL((k_outer_finish<K>), dim3(1), dim3(kBlock), 0, s, pr, gs, ost, run_mask, inv_scale, tol, 0);
    CC_CHECK(hipGetLastError());
  }
  CC_CHECK(hipMemcpyAsync(ho.data(), ost, K * sizeof(OuterState), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  for (int c = 0; c < K; ++c) {
    results[c].iterations = inner_total[c];
    results[c].status = ho[c].status == CC_CONVERGED ? CC_CONVERGED : CC_MAXITER;
    results[c].outer = outer[c];
    results[c].inner_total = inner_total[c];
    results[c].recursive_rel = ho[c].rel;
    results[c].true_rel = ho[c].rel;
  }
  return 0;
}

int refine_operator_multi(int n, int k, const std::function<int(const double*, double*, double*, int*)>& residual,
                          const std::function<int(const double*, double*, int, int*)>& inner, double* X, double tol,
                          int maxiter, int max_outer, void* ws, size_t wsb, cc_pcg_result* results, hipStream_t s) {
  switch (k) {
    case 1: return refine_operator_k<1>(n, residual, inner, X, tol, maxiter, max_outer, ws, wsb, results, s);
    case 2: return refine_operator_k<2>(n, residual, inner, X, tol, maxiter, max_outer, ws, wsb, results, s);
    case 4: return refine_operator_k<4>(n, residual, inner, X, tol, maxiter, max_outer, ws, wsb, results, s);
    case 8: return refine_operator_k<8>(n, residual, inner, X, tol, maxiter, max_outer, ws, wsb, results, s);
    case 16: return refine_operator_k<16>(n, residual, inner, X, tol, maxiter, max_outer, ws, wsb, results, s);
    default: return CC_E_ARG;
  }
}

size_t refine_operator_multi_bytes(size_t n, int k) {
  switch (k) {
    case 1: return refine_bytes<1>(n);
    case 2: return refine_bytes<2>(n);
    case 4: return refine_bytes<4>(n);
    case 8: return refine_bytes<8>(n);
    case 16: return refine_bytes<16>(n);
    default: return 0;
  }
}

int cg_operator_mixed_multi(int n, int k, const std::function<int(const double*, double*, double*, int*)>& residual,
                            const std::function<int(const float*, float*, double*, int*)>& apply32,
                            const double* diag_src, double* X, double tol, int maxiter, double inner_tol,
                            int max_outer, int precond, int check_every, void* ws, size_t wsb,
                            cc_pcg_result* results, hipStream_t s) {
#define CC_MXK(KK)                                                                                          \
  case KK:                                                                                                  \
    return mixed_operator_k<KK>(n, residual, apply32, diag_src, X, tol, maxiter, inner_tol, max_outer, precond, \
                                check_every, ws, wsb, results, s);
  switch (k) {
    CC_MXK(1)
    CC_MXK(2)
    CC_MXK(4)
    CC_MXK(8)
    CC_MXK(16)
    default: return CC_E_ARG;
  }
#undef CC_MXK
}

size_t cg_operator_mixed_multi_bytes(size_t n, int k) {
  switch (k) {
    case 1: return mixed_bytes<1>(n);
    case 2: return mixed_bytes<2>(n);
    case 4: return mixed_bytes<4>(n);
    case 8: return mixed_bytes<8>(n);
    case 16: return mixed_bytes<16>(n);
    default: return 0;
  }
}

// ---------------------------------------------------------------------------
// Fixed device scratch for cc_pcg_jacobi_f64_noalloc.
//
// One file-scope array per GPU (the code object's global lives on every
// device the library runs on, until the process ends). Host code never uses
// the array's name as an address: at a device's first call the address is
// read once with hipGetSymbolAddress, with that device current, and kept per
// device. The solver writes every word of the workspace it reads before it
// reads it (k_cg_setup, k_diag and the producer kernels), so no clearing is
// needed between calls; calls that overlap on one GPU share the array and
// race.
// ---------------------------------------------------------------------------

constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static void* g_scratch_ptr[kScratchDevices];
static int g_scratch_status[kScratchDevices];
static std::once_flag g_scratch_once[kScratchDevices];

static int scratch_for_current_device(void** out) {
  int dev = -1;
  CC_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) return CC_E_ARG;
  std::call_once(g_scratch_once[dev], [dev] {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_status[dev] = (int)e;
    g_scratch_ptr[dev] = e == hipSuccess ? p : nullptr;
  });
  if (g_scratch_status[dev] != 0) return g_scratch_status[dev];
  if (!g_scratch_ptr[dev]) return CC_E_NODEVICE;
  *out = g_scratch_ptr[dev];
  return 0;
}

// Hooks used by the batched solver (batched.hip).
void prof_begin_batched(hipStream_t s) { prof_begin(kProfBatched, s); }
void prof_end_batched(hipStream_t s) { prof_end(s); }

}  // namespace cc

using namespace cc;

// ---------------------------------------------------------------------------
// C ABI.
// ---------------------------------------------------------------------------

#define CC_DISPATCH_K(k, CALL)        \
  switch (k) {                        \
    case 1: { constexpr int KK = 1; return CALL; }  \
    case 2: { constexpr int KK = 2; return CALL; }  \
    case 4: { constexpr int KK = 4; return CALL; }  \
    case 8: { constexpr int KK = 8; return CALL; }  \
    case 16: { constexpr int KK = 16; return CALL; } \
    default: return CC_E_ARG;         \
  }

extern "C" {

int cc_version(void) { return 10100; }

int cc_device_count(int* count) {
  if (!count) return CC_E_ARG;
  return (int)hipGetDeviceCount(count);
}

int cc_profile_enable_classes(int32_t mask) {
  g_prof.classes = (uint32_t)mask & ((1u << kProfClasses) - 1u);
  g_prof.on = g_prof.classes != 0;
  g_prof.open = false;
  g_prof.pending.clear();
  g_prof.used = 0;
  return 0;
}

int cc_profile_enable(int32_t on) { return cc_profile_enable_classes(on ? -1 : 0); }

int cc_profile_reset(void) {
  for (int i = 0; i < kProfClasses; ++i) {
    g_prof.ms[i] = 0.0;
    g_prof.count[i] = 0;
  }
  return 0;
}

int cc_profile_read(int32_t which, double* total_ms, int64_t* count) {
  if (which < 0 || which >= kProfClasses || !total_ms || !count) return CC_E_ARG;
  *total_ms = g_prof.ms[which];
  *count = g_prof.count[which];
  return 0;
}

int cc_profile_collect(void) {
  prof_collect();
  return 0;
}

int cc_spmm_set_config(int32_t k, int32_t value_bytes, int32_t nz_lanes, int32_t grid, int32_t unroll,
                       int32_t cols_per_lane, int32_t stream_loads, int32_t sweep_steps) {
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16) return CC_E_ARG;
  if (value_bytes != 4 && value_bytes != 8) return CC_E_ARG;
  if (cols_per_lane * value_bytes > 16) return CC_E_ARG;
  if (cols_per_lane != 0 && cols_per_lane != 1 && cols_per_lane != 2 && cols_per_lane != 4) return CC_E_ARG;
  if (cols_per_lane > k) return CC_E_ARG;
  const int kl = k / (cols_per_lane > 0 ? cols_per_lane : 1);
  if (nz_lanes < 0 || kl * nz_lanes > 64 || (nz_lanes & (nz_lanes - 1)) != 0) return CC_E_ARG;
  if (grid < 0 || grid > kMaxParts) return CC_E_ARG;
  if (unroll != 0 && unroll != 1 && unroll != 4 && unroll != 8 && !(unroll == kUnrollWindow && k == 1)) return CC_E_ARG;
  if (stream_loads != 0 && (stream_loads != 1 || unroll == 1)) return CC_E_ARG;
  if (sweep_steps < 0 || sweep_steps > kMaxSweep) return CC_E_ARG;
  g_spmm_cfg[value_bytes == 4 ? 1 : 0][k] = SpmmConfig{nz_lanes, grid, unroll, cols_per_lane, stream_loads, sweep_steps};
  return 0;
}

int cc_spmv_csr_f64(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals, const double* x,
                    double* y, cc_stream_t stream) {
  if (n < 0 || (n > 0 && (!indptr || !indices || !vals || !x || !y))) return CC_E_ARG;
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  return spmm_launch<double, double, 1, 0>(n, avg_row_if_needed<double>(n, 1, indptr, s), indptr, indices, vals, x,
                                           y, nullptr, nullptr, nullptr, s, nullptr);
}

int cc_spmv_csr_f32(int32_t n, const int32_t* indptr, const int32_t* indices, const float* vals, const float* x,
                    float* y, cc_stream_t stream) {
  if (n < 0 || (n > 0 && (!indptr || !indices || !vals || !x || !y))) return CC_E_ARG;
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  return spmm_launch<float, float, 1, 0>(n, avg_row_if_needed<float>(n, 1, indptr, s), indptr, indices, vals, x, y,
                                         nullptr, nullptr, nullptr, s, nullptr);
}

int cc_spmm_csr_f64(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                    const double* X, double* Y, cc_stream_t stream) {
  if (n < 0 || (n > 0 && (!indptr || !indices || !vals || !X || !Y))) return CC_E_ARG;
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16) return CC_E_ARG;
  const double avg = avg_row_if_needed<double>(n, k, indptr, s);
  CC_DISPATCH_K(k, (spmm_launch<double, double, KK, 0>(n, avg, indptr, indices, vals, X, Y, nullptr, nullptr, nullptr,
                                                       s, nullptr)));
}

int cc_spmm_csr_f32(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const float* vals,
                    const float* X, float* Y, cc_stream_t stream) {
  if (n < 0 || (n > 0 && (!indptr || !indices || !vals || !X || !Y))) return CC_E_ARG;
  if (n == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16) return CC_E_ARG;
  const double avg = avg_row_if_needed<float>(n, k, indptr, s);
  CC_DISPATCH_K(k, (spmm_launch<float, float, KK, 0>(n, avg, indptr, indices, vals, X, Y, nullptr, nullptr, nullptr,
                                                     s, nullptr)));
}

int cc_jacobi_diag_f64(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals, double* diag,
                       cc_stream_t stream) {
  if (n < 0 || (n > 0 && (!indptr || !indices || !vals || !diag))) return CC_E_ARG;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_diag, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, indptr, indices, vals,
                     diag);
  return (int)hipGetLastError();
}

int cc_residual_f64(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                    const double* B, const double* X, double* rel, void* workspace, size_t workspace_bytes,
                    cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !vals || !B || !X || !rel || !workspace) return CC_E_ARG;
  const size_t need = (size_t)n * k * sizeof(double) + (size_t)kMaxParts * 2 * k * sizeof(double) + 512;
  if (workspace_bytes < need) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  Carve cv{(char*)workspace, 0, workspace_bytes};
  double* R = cv.take<double>((size_t)n * k);
  double* part = cv.take<double>((size_t)kMaxParts * 2 * k);
  const double avg = avg_row_len(n, indptr, s);
  CC_DISPATCH_K(k, (true_residual<KK>(n, avg, indptr, indices, vals, B, X, R, part, s, rel)));
}

size_t cc_pcg_workspace_bytes(int32_t n, int32_t k, int32_t precision) {
  if (n <= 0 || k <= 0) return 0;
  if (precision == 1) {
    switch (k) {
      case 1: return mixed_bytes<1>(n);
      case 2: return mixed_bytes<2>(n);
      case 4: return mixed_bytes<4>(n);
      case 8: return mixed_bytes<8>(n);
      case 16: return mixed_bytes<16>(n);
      default: return 0;
    }
  }
  return cg_bytes<double>(n, k);
}

int cc_pcg_jacobi_f64(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                      const double* B, double* X, double tol, int32_t maxiter, int32_t preconditioner,
                      int32_t check_every, void* workspace, size_t workspace_bytes, cc_pcg_result* results,
                      double* history, cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !vals || !B || !X || !results || !workspace || maxiter < 0) return CC_E_ARG;
  if (workspace_bytes < cc_pcg_workspace_bytes(n, k, 0)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  CC_DISPATCH_K(k, (pcg_f64<KK>(n, indptr, indices, vals, B, X, tol, maxiter, preconditioner, check_every, workspace,
                                workspace_bytes, results, history, s)));
}

int cc_pcg_jacobi_mixed(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                        const float* vals32, const double* B, double* X, double tol, int32_t maxiter,
                        double inner_tol, int32_t max_outer, int32_t preconditioner, int32_t check_every,
                        void* workspace, size_t workspace_bytes, cc_pcg_result* results, cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !vals || !vals32 || !B || !X || !results || !workspace || maxiter < 0)
    return CC_E_ARG;
  if (workspace_bytes < cc_pcg_workspace_bytes(n, k, 1)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  CC_DISPATCH_K(k, (pcg_mixed<KK>(n, indptr, indices, vals, vals32, B, X, tol, maxiter, inner_tol, max_outer,
                                  preconditioner, check_every, workspace, workspace_bytes, results, s)));
}

int cc_pcg_jacobi_f64_rb(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                         const cc_rb_desc* rb, const double* B, double* X, double tol, int32_t maxiter,
                         int32_t preconditioner, int32_t check_every, void* workspace, size_t workspace_bytes,
                         cc_pcg_result* results, double* history, cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !vals || !B || !X || !results || !workspace || maxiter < 0) return CC_E_ARG;
  if (!rb_desc_ok(rb, n) || !rb->gval) return CC_E_ARG;
  if (workspace_bytes < cc_pcg_workspace_bytes(n, k, 0)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  CC_DISPATCH_K(k, (pcg_f64<KK>(n, indptr, indices, vals, B, X, tol, maxiter, preconditioner, check_every, workspace,
                                workspace_bytes, results, history, s, rb)));
}

int cc_pcg_jacobi_mixed_rb(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                           const cc_rb_desc* rb, const double* B, double* X, double tol, int32_t maxiter,
                           double inner_tol, int32_t max_outer, int32_t preconditioner, int32_t check_every,
                           void* workspace, size_t workspace_bytes, cc_pcg_result* results, cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !vals || !B || !X || !results || !workspace || maxiter < 0) return CC_E_ARG;
  if (!rb_desc_ok(rb, n) || !rb->gval32) return CC_E_ARG;
  if (workspace_bytes < cc_pcg_workspace_bytes(n, k, 1)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  // The inner solve never reads the CSR float32 values when rb is given.
  CC_DISPATCH_K(k, (pcg_mixed<KK>(n, indptr, indices, vals, nullptr, B, X, tol, maxiter, inner_tol, max_outer,
                                  preconditioner, check_every, workspace, workspace_bytes, results, s, rb)));
}


int cc_pcg_jacobi_f64_runs(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                           const int32_t* rptr, const uint32_t* runs, const double* B, double* X, double tol,
                           int32_t maxiter, int32_t preconditioner, int32_t check_every, void* workspace,
                           size_t workspace_bytes, cc_pcg_result* results, double* history, cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !vals || !rptr || !runs || !B || !X || !results || !workspace || maxiter < 0)
    return CC_E_ARG;
  if (workspace_bytes < cc_pcg_workspace_bytes(n, 1, 0)) return CC_E_WORKSPACE;
  const RunsView rv{rptr, runs};
  return pcg_f64<1>(n, indptr, indices, vals, B, X, tol, maxiter, preconditioner, check_every, workspace,
                    workspace_bytes, results, history, (hipStream_t)stream, nullptr, &rv);
}

int cc_pcg_jacobi_mixed_runs(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                             const float* vals32, const int32_t* rptr, const uint32_t* runs, const double* B,
                             double* X, double tol, int32_t maxiter, double inner_tol, int32_t max_outer,
                             int32_t preconditioner, int32_t check_every, void* workspace, size_t workspace_bytes,
                             cc_pcg_result* results, cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !vals || !vals32 || !rptr || !runs || !B || !X || !results || !workspace ||
      maxiter < 0)
    return CC_E_ARG;
  if (workspace_bytes < cc_pcg_workspace_bytes(n, 1, 1)) return CC_E_WORKSPACE;
  const RunsView rv{rptr, runs};
  return pcg_mixed<1>(n, indptr, indices, vals, vals32, B, X, tol, maxiter, inner_tol, max_outer, preconditioner,
                      check_every, workspace, workspace_bytes, results, (hipStream_t)stream, nullptr, &rv);
}

size_t cc_pcg_noalloc_max_bytes(void) { return kScratchBytes; }

int cc_pcg_jacobi_f64_noalloc(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices,
                              const double* vals, const double* B, double* X, double tol, int32_t maxiter,
                              int32_t preconditioner, int32_t check_every, cc_pcg_result* results, double* history,
                              cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !vals || !B || !X || !results || maxiter < 0) return CC_E_ARG;
  if (k != 1 && k != 2 && k != 4 && k != 8 && k != 16) return CC_E_ARG;
  const size_t need = cc_pcg_workspace_bytes(n, k, 0);
  if (need == 0 || need > kScratchBytes) return CC_E_TOO_LARGE;
  void* ws = nullptr;
  CC_CHECK(scratch_for_current_device(&ws));
  hipStream_t s = (hipStream_t)stream;
  CC_DISPATCH_K(k, (pcg_f64<KK>(n, indptr, indices, vals, B, X, tol, maxiter, preconditioner, check_every, ws,
                                kScratchBytes, results, history, s)));
}

}  // extern "C"
