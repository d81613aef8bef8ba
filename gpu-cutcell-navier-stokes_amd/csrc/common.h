// common.h - shared device-side types of the cut-cell PCG library.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cc {

// Upper bound on the per-block partial sums a producer kernel emits; the
// consumer kernel re-reduces them in a fixed order.
constexpr int kMaxParts = 1024;

// Kernel classes of the optional event timing (cc_profile_*).
constexpr int kProfSpmm = 0;
constexpr int kProfUpdate = 1;
constexpr int kProfDirection = 2;
constexpr int kProfBatched = 3;
constexpr int kProfLearned = 4;
constexpr int kProfClasses = 8;

// Per right-hand-side state of one CG recurrence (device memory). Only the
// one-block finalize kernels write it; the streaming kernels read it.
struct ColState {
  double rho;
  double bnorm;
  double rel;
  double alpha;
  double beta;
  int32_t status;
  int32_t iterations;
};

// Per right-hand-side state of the outer (float64) refinement loop.
struct OuterState {
  double bnorm;
  double rel;
  double scale;
  double best_rel;
  int32_t status;
  int32_t outer;
  int32_t inner_total;
  int32_t pad;
};

// Event-timing hooks shared by the solver sources (defined in pcg.hip).
void prof_begin_batched(hipStream_t s);
void prof_end_batched(hipStream_t s);
void prof_begin(int cls, hipStream_t s);
void prof_end(hipStream_t s);
void prof_collect();
// Timing of a step that is exactly one kernel launch: prof_arm(cls) before
// it, prof_disarm() after. The launcher takes the class's event pair with
// prof_take() and hands it to hipExtLaunchKernelGGL, which records the pair
// with the kernel's own dispatch rather than as two separate stream
// commands. Timing every product of the 16-column mixed solve this way slows
// the solve by about 2%; two separate event records per product cost about 4%.
void prof_arm(int cls);
void prof_disarm();
void prof_take(hipEvent_t* start, hipEvent_t* stop);

// hipLaunchKernelGGL that carries the armed event pair, if any.
template <typename F, typename... Args>
inline void launch_timed(F kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t s, Args... args) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  prof_take(&e0, &e1);
  hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, e0, e1, 0, args...);
}

template <typename T>
struct Accum {
  using type = T;
};

// One axis of a grid coarsening (mg.hip): fine and coarse point counts and
// whether the axis is periodic.
struct MgAxis {
  int nf, nc, per;
};
struct MgDims {
  MgAxis a[3];
};

// Prolongation sources of fine point f: count 1 or 2, coarse indices and weights.
__device__ inline int pro_src(int f, const MgAxis& ax, int* c, float* w) {
  if ((f & 1) == 0) {
    c[0] = f >> 1;
    w[0] = 1.0f;
    return 1;
  }
  c[0] = (f - 1) >> 1;
  w[0] = 0.5f;
  const int c1 = c[0] + 1;
  if (c1 < ax.nc) {
    c[1] = c1;
    w[1] = 0.5f;
    return 2;
  }
  if (ax.per) {
    c[1] = 0;
    w[1] = 0.5f;
    return 2;
  }
  w[0] = 1.0f;
  return 1;
}

// The 8 prolongation taps of fine point (fx, fy, fz): coarse point offsets
// (x fastest) and weights wz * wy * wx. A one-source axis repeats its source
// at weight 0, so the 8 loads of a sum are independent; adding 0 * c leaves
// every sum as the taps-only loop has it. The weights are 0, 1/8 ... 1, so
// their products are exact.
__device__ inline void prolong_taps(const MgDims& g, int fx, int fy, int fz, size_t* off, float* w) {
  int cx[2], cy[2], cz[2];
  float wx[2], wy[2], wz[2];
  if (pro_src(fx, g.a[0], cx, wx) == 1) {
    cx[1] = cx[0];
    wx[1] = 0.0f;
  }
  if (pro_src(fy, g.a[1], cy, wy) == 1) {
    cy[1] = cy[0];
    wy[1] = 0.0f;
  }
  if (pro_src(fz, g.a[2], cz, wz) == 1) {
    cz[1] = cz[0];
    wz[1] = 0.0f;
  }
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int k = t >> 2, j = (t >> 1) & 1, i = t & 1;
    off[t] = ((size_t)cz[k] * g.a[1].nc + cy[j]) * g.a[0].nc + cx[i];
    w[t] = wz[k] * wy[j] * wx[i];
  }
}

// acc + (P coarse)(fx, fy, fz), coarse a single-channel grid with element
// stride cs: the 8 loads issued together, then the terms added in tap order.
// Every kernel that prolongates sums this way, so the results agree bit for
// bit.
__device__ inline float prolong_value(const MgDims& g, const float* __restrict__ coarse, int cs, float acc, int fx,
                                      int fy, int fz) {
  size_t off[8];
  float w[8], cv[8];
  prolong_taps(g, fx, fy, fz, off, w);
#pragma unroll
  for (int t = 0; t < 8; ++t) cv[t] = coarse[off[t] * cs];
#pragma unroll
  for (int t = 0; t < 8; ++t) acc += w[t] * cv[t];
  return acc;
}

}  // namespace cc

#include <functional>

struct cc_pcg_result;

namespace cc {
// Float64 single-right-hand-side Jacobi-PCG over an operator (pcg.hip).
// apply(P, Q, partial, &nparts, state) launches Q = A P with per-block
// partials of P.Q; residual(X, scratch, &rel) computes ||B - A X|| / ||B||.
// snapshots[j] (device, [n_snap][n]) receives the iterate after
// snap_iters[j] (host) iterations.
int cg_operator_f64(int n, const std::function<int(const double*, double*, double*, int*, const ColState*)>& apply,
                    const double* diag_src, const double* B, double* X, double tol, int maxiter, int precond,
                    int check_every, void* ws, size_t wsb, cc_pcg_result* result, double* hist, hipStream_t s,
                    const std::function<int(const double*, double*, double*)>& residual,
                    const int* snap_iters = nullptr, int n_snap = 0, double* snapshots = nullptr);
size_t cg_operator_bytes(size_t n);
// The same for k in {1, 2, 4, 8, 16} right-hand sides (row-major [n][k]);
// apply writes per-block, per-column partials [nparts][k]; residual writes k
// relative residuals to host memory.
// tail: rows after P and Q ([n + tail][k]) that the operator may use.
int cg_operator_multi(int n, int k, const std::function<int(const double*, double*, double*, int*)>& apply,
                      const double* diag_src, const double* B, double* X, double tol, int maxiter, int precond,
                      int check_every, void* ws, size_t wsb, cc_pcg_result* results, hipStream_t s,
                      const std::function<int(const double*, double*, double*)>& residual, size_t tail = 0);
size_t cg_operator_multi_bytes(size_t n, int k, size_t tail = 0);
// FP32 inner / FP64 outer refinement over an operator for k in {1, 2, 4, 8,
// 16} columns: residual(X, R64, partial, &nparts) writes R64 = B - A X with
// per-block, per-column partials of R64.R64 (at most kMaxParts blocks);
// apply32(P, Q, partial, &nparts) is the float32 operator with per-block
// partials of P.Q. X is zeroed first (pcg.hip).
int cg_operator_mixed_multi(int n, int k, const std::function<int(const double*, double*, double*, int*)>& residual,
                            const std::function<int(const float*, float*, double*, int*)>& apply32,
                            const double* diag_src, double* X, double tol, int maxiter, double inner_tol,
                            int max_outer, int precond, int check_every, void* ws, size_t wsb,
                            cc_pcg_result* results, hipStream_t s);
size_t cg_operator_mixed_multi_bytes(size_t n, int k);
// Iterative refinement around an inner solver (k in {1, 2, 4, 8, 16}):
// residual as above; inner(Bs, Z, maxiter_left, its) solves for Z from the
// scaled float64 residual Bs (columns not running have Bs = 0) and writes
// each column's iteration count to its (host). X is zeroed first (pcg.hip).
int refine_operator_multi(int n, int k, const std::function<int(const double*, double*, double*, int*)>& residual,
                          const std::function<int(const double*, double*, int, int*)>& inner, double* X, double tol,
                          int maxiter, int max_outer, void* ws, size_t wsb, cc_pcg_result* results, hipStream_t s);
size_t refine_operator_multi_bytes(size_t n, int k);
// Y = A X (CSR, float64, k in {1, 2, 4, 8, 16}) with the autotuned launch
// shape; partial[block][c] gets per-block X.Y sums when not null (pcg.hip).
int csr_spmm_f64(int n, int k, double avg_row, const int* indptr, const int* indices, const double* vals,
                 const double* X, double* Y, double* partial, hipStream_t s, int* gs);
}  // namespace cc

struct cc_cnn_desc;
namespace cc {
// Flexible CG with the learned preconditioner over an operator (fcg.hip).
int fcg_operator(int n, const std::function<int(const double*, double*, double*, int*)>& apply,
                 const cc_cnn_desc& net, const double* diag, const double* bvec, double* x, double tol, int maxiter,
                 int check_every, void* ws, size_t wsb, cc_pcg_result* result, double* hist, hipStream_t s,
                 const std::function<int(const double*, double*, double*)>& residual);
size_t fcg_operator_bytes(size_t n, size_t voxels);
}  // namespace cc

struct cc_mlcnn_desc;
namespace cc {
// Flexible CG with the multilevel learned preconditioner (fcg.hip).
int fcg_operator_ml(int n, const std::function<int(const double*, double*, double*, int*)>& apply,
                    const cc_mlcnn_desc& ml, const double* diag, const double* bvec, double* x, double tol,
                    int maxiter, int check_every, void* ws, size_t wsb, cc_pcg_result* result, double* hist,
                    hipStream_t s, const std::function<int(const double*, double*, double*)>& residual);
size_t fcg_operator_ml_bytes(size_t n, const cc_mlcnn_desc& ml);
bool ml_desc_ok(const cc_mlcnn_desc& ml);
// The same on k in {1, 2, 4, 8, 16} right-hand sides (row-major [n][k]);
// apply writes per-block, per-column partials [nparts][k] of p.q; residual
// writes k relative residuals to host memory.
int fcg_operator_ml_k(int n, int k, const std::function<int(const double*, double*, double*, int*)>& apply,
                      const cc_mlcnn_desc& ml, const double* diag, const double* bvec, double* x, double tol,
                      int maxiter, int check_every, void* ws, size_t wsb, cc_pcg_result* results, hipStream_t s,
                      const std::function<int(const double*, double*, double*)>& residual);
size_t fcg_operator_ml_k_bytes(size_t n, const cc_mlcnn_desc& ml, int k);
}  // namespace cc

struct cc_rb_desc;
namespace cc {
// Row-block product Y = A X (rowblock.hip); partial[block][c] gets X.Y sums
// when not NULL. With B (not for the paired form): the residual Y = B - A X
// and partials Y.Y. MT = VT = double or float; K in {1, 2, 4, 8, 16}.
template <typename MT, typename VT, int K>
int rb_spmm(const cc_rb_desc& rb, const VT* X, VT* Y, double* partial, const ColState* st, hipStream_t s,
            int* grid_out, const VT* B = nullptr);
bool rb_desc_ok(const cc_rb_desc* rb, int n);
}  // namespace cc

#define CC_CHECK(expr)             \
  do {                             \
    int cc_err_ = (int)(expr);     \
    if (cc_err_ != 0) return cc_err_; \
  } while (0)
