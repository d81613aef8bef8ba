// common.h - shared device-side types of the cut-cell PCG library.
#pragma once

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

template <typename T>
struct Accum {
  using type = T;
};

}  // namespace cc

#define CC_CHECK(expr)             \
  do {                             \
    int cc_err_ = (int)(expr);     \
    if (cc_err_ != 0) return cc_err_; \
  } while (0)
