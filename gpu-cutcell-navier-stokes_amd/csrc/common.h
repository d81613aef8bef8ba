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
// when not NULL. MT = VT = double or float; K in {1, 2, 4, 8, 16}.
template <typename MT, typename VT, int K>
int rb_spmm(const cc_rb_desc& rb, const VT* X, VT* Y, double* partial, const ColState* st, hipStream_t s,
            int* grid_out);
bool rb_desc_ok(const cc_rb_desc* rb, int n);
}  // namespace cc

#define CC_CHECK(expr)             \
  do {                             \
    int cc_err_ = (int)(expr);     \
    if (cc_err_ != 0) return cc_err_; \
  } while (0)
