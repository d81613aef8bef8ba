// ekte.hip - the aggregated operator y = E^T B (E x) applied without forming
// E^T B E, for k in {1, 2, 4, 8, 16} right-hand sides (row-major [n][k]).
//
// E maps the reduced unknowns (the interior dofs) to the active background
// dofs: an interior dof's row of E is the unit vector of its reduced index, a
// small dof's row holds its polynomial-preserving extension weights
// (code/bell_3d_core.py:471-542). The reference forms E^T A E with a sparse
// triple product (restrict_spd, code/bell_3d_core.py:550-552) and multiplies
// by it (code/fast_solver_backends.py:250). Here the gather and the scatter of
// E are folded into the product on the background operator, over a combined
// vector whose first n_red rows are x and whose last n_small rows hold the
// small dofs' values:
//   G  u_s = E_s x                       small dofs only (one short row each),
//                                        written after x
//   S  [y ; v_s] = B [x ; u_s]           B over the active dofs, its rows and
//                                        columns renumbered at set-up so that
//                                        interior dof j is index j and small
//                                        dof s is index n_red + s; the
//                                        autotuned CSR (or row-block) product
//   T  y_j += sum_s E[s, j] v_s          for the reduced unknowns that some
//                                        small dof's stencil reaches, each
//                                        summed in a fixed order by one lane
// No full-grid vector is materialised for the interior dofs and nothing is
// summed with atomics. Inside CG the combined vectors are P and Q themselves
// (the CG buffers carry n_small tail rows), and S's per-block x.y partials
// over all combined rows are exactly p.(E^T B E)p = (E p).B(E p).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/cutcell_hip.h"
#include "common.h"
#include "spmm_util.h"

namespace cc {

namespace {

// One wave per short sparse row r of W (out = W v, K columns): lane (z, c)
// owns CPL consecutive columns c (one 8- or 16-byte load) and the row's
// entries z, z + NZ, ... A chunk of up to 8 entries per lane is handled in
// two memory rounds: all of its indices and weights, then all of its
// gathered v rows. The NZ slot sums of a column are combined by a fixed
// shuffle tree. ADD: out += sum.
template <int K, bool ADD>
__device__ inline void wave_row_product(int r, const int* __restrict__ w_ptr, const int* __restrict__ w_col,
                                        const double* __restrict__ w_val, const double* __restrict__ v,
                                        double* __restrict__ out, int out_row) {
  constexpr int CPL = K >= 2 ? 2 : 1;
  constexpr int KL = K / CPL;
  constexpr int NZ = 64 / KL;
  constexpr int U = 8;
  const int lane = threadIdx.x & 63;
  const int c0 = (lane % KL) * CPL;
  const int z = lane / KL;
  const int e0 = w_ptr[r], e1 = w_ptr[r + 1];
  double acc[CPL];
#pragma unroll
  for (int m = 0; m < CPL; ++m) acc[m] = 0.0;
  for (int base = e0 + z; base < e1; base += U * NZ) {
    int q[U];
    double w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = base + u * NZ;
      const bool ok = e < e1;
      q[u] = ok ? w_col[e] : 0;
      w[u] = ok ? w_val[e] : 0.0;
    }
    double xv[U][CPL];
#pragma unroll
    for (int u = 0; u < U; ++u) vload<double, CPL>(v + (size_t)q[u] * K + c0, xv[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int m = 0; m < CPL; ++m) acc[m] += w[u] * xv[u][m];
  }
#pragma unroll
  for (int m = 0; m < CPL; ++m)
#pragma unroll
    for (int o = KL; o < 64; o <<= 1) acc[m] += __shfl_xor(acc[m], o, 64);
  if (z == 0) {
    double* dst = out + (size_t)out_row * K + c0;
    if constexpr (ADD) {
      double cur[CPL];
      vload<double, CPL>(dst, cur);
#pragma unroll
      for (int m = 0; m < CPL; ++m) acc[m] += cur[m];
    }
    vstore<double, CPL>(dst, acc);
  }
}

// G: u[s] = E_s[s, :] x, one wave per small dof (four waves per block).
template <int K>
__global__ __launch_bounds__(256) void k_ekte_gather(int n_small, const int* __restrict__ e_ptr,
                                                     const int* __restrict__ e_col, const double* __restrict__ e_w,
                                                     const double* __restrict__ X, double* __restrict__ U) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n_small) return;
  wave_row_product<K, false>(r, e_ptr, e_col, e_w, X, U, r);
}

// T: y[rows[i]] += T[i, :] v, one wave per touched unknown, the row's
// entries summed in a fixed order.
template <int K>
__global__ __launch_bounds__(256) void k_ekte_touch(int n_touch, const int* __restrict__ rows,
                                                    const int* __restrict__ t_ptr, const int* __restrict__ t_small,
                                                    const double* __restrict__ t_w, const double* __restrict__ V,
                                                    double* __restrict__ Y) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n_touch) return;
  wave_row_product<K, true>(i, t_ptr, t_small, t_w, V, Y, rows[i]);
}

template <int K>
__global__ __launch_bounds__(256) void k_ekte_residual(int n, const double* __restrict__ B,
                                                       const double* __restrict__ AX, double* __restrict__ part) {
  __shared__ double red[4][2 * K];
  const long long total = (long long)n * K;
  double rr = 0.0, bb = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const double bv = B[t];
    const double r = bv - AX[t];
    rr += r * r;
    bb += bv * bv;
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = K; o < 64; o <<= 1) {
    rr += __shfl_xor(rr, o, 64);
    bb += __shfl_xor(bb, o, 64);
  }
  if (lane < K) {
    red[wave][lane] = rr;
    red[wave][K + lane] = bb;
  }
  __syncthreads();
  if (threadIdx.x < 2 * K) {
    const int c = threadIdx.x;
    part[(size_t)blockIdx.x * 2 * K + c] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
  }
}

bool ekte_ok(const cc_ekte_desc* d) {
  if (!d || d->n_red <= 0 || d->n_small < 0 || d->n_touch < 0 || d->nnz_b < 0) return false;
  if (!d->b_indptr || !d->b_indices || !d->b_vals) return false;
  if (d->n_small > 0 && (!d->e_indptr || !d->e_indices || !d->e_vals)) return false;
  if (d->n_touch > 0 && (!d->t_rows || !d->t_indptr || !d->t_small || !d->t_vals)) return false;
  if (d->b_rb && (!rb_desc_ok(d->b_rb, d->n_red + d->n_small) || !d->b_rb->gval)) return false;
  return true;
}

bool k_ok(int k) { return k == 1 || k == 2 || k == 4 || k == 8 || k == 16; }

int grid_for(long long work, int cap) {
  long long g = (work + 255) / 256;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

size_t round_up(size_t b) { return (b + 255) & ~size_t(255); }

// Two combined vectors [n_red + n_small][k] (the standalone apply's copies).
size_t apply_bytes(const cc_ekte_desc& d, int k) {
  return 2 * round_up((size_t)(d.n_red + d.n_small) * k * sizeof(double)) + 256;
}

// The three stages on combined vectors: XU's first n_red rows are the input
// (its tail is overwritten with u_s), YV receives [y ; v_s] and then the
// scatter into y. partial (may be null) receives S's per-block x.y sums.
template <int K>
int ekte_apply_k(const cc_ekte_desc& d, double* XU, double* YV, double* partial, int* gs, hipStream_t s) {
  const int rows = d.n_red + d.n_small;
  if (d.n_small > 0) {
    hipLaunchKernelGGL((k_ekte_gather<K>), dim3((unsigned)((d.n_small + 3) / 4)), dim3(256), 0, s, d.n_small,
                       d.e_indptr, d.e_indices, d.e_vals, XU, XU + (size_t)d.n_red * K);
    CC_CHECK(hipGetLastError());
  }
  if (d.b_rb) {
    CC_CHECK((rb_spmm<double, double, K>(*d.b_rb, XU, YV, partial, nullptr, s, gs)));
  } else {
    const double avg = rows > 0 ? (double)d.nnz_b / rows : 0.0;
    CC_CHECK(csr_spmm_f64(rows, K, avg, d.b_indptr, d.b_indices, d.b_vals, XU, YV, partial, s, gs));
  }
  if (d.n_touch > 0) {
    hipLaunchKernelGGL((k_ekte_touch<K>), dim3((unsigned)((d.n_touch + 3) / 4)), dim3(256), 0, s, d.n_touch, d.t_rows,
                       d.t_indptr, d.t_small, d.t_vals, YV + (size_t)d.n_red * K, YV);
  }
  return (int)hipGetLastError();
}

int ekte_apply(const cc_ekte_desc& d, int k, double* XU, double* YV, double* partial, int* gs, hipStream_t s) {
  switch (k) {
    case 1: return ekte_apply_k<1>(d, XU, YV, partial, gs, s);
    case 2: return ekte_apply_k<2>(d, XU, YV, partial, gs, s);
    case 4: return ekte_apply_k<4>(d, XU, YV, partial, gs, s);
    case 8: return ekte_apply_k<8>(d, XU, YV, partial, gs, s);
    case 16: return ekte_apply_k<16>(d, XU, YV, partial, gs, s);
    default: return CC_E_ARG;
  }
}

// Y = E^T B E X for caller vectors of n_red rows: X is copied into the
// combined buffer XU and y copied out of YV.
int ekte_apply_plain(const cc_ekte_desc& d, int k, const double* X, double* Y, double* XU, double* YV,
                     hipStream_t s) {
  const size_t bytes = (size_t)d.n_red * k * sizeof(double);
  CC_CHECK(hipMemcpyAsync(XU, X, bytes, hipMemcpyDeviceToDevice, s));
  CC_CHECK(ekte_apply(d, k, XU, YV, nullptr, nullptr, s));
  return (int)hipMemcpyAsync(Y, YV, bytes, hipMemcpyDeviceToDevice, s);
}

// ||B - A X|| / ||B|| per column (||B|| = 0 -> 1), summed on the host in
// block order; scratch holds n_red rows.
int ekte_true_residual(const cc_ekte_desc& d, int k, const double* B, const double* X, double* scratch, double* XU,
                       double* YV, double* part, double* rel, hipStream_t s) {
  CC_CHECK(ekte_apply_plain(d, k, X, scratch, XU, YV, s));
  const int g = grid_for((long long)d.n_red * k, 256);
  switch (k) {
    case 1: hipLaunchKernelGGL((k_ekte_residual<1>), dim3(g), dim3(256), 0, s, d.n_red, B, scratch, part); break;
    case 2: hipLaunchKernelGGL((k_ekte_residual<2>), dim3(g), dim3(256), 0, s, d.n_red, B, scratch, part); break;
    case 4: hipLaunchKernelGGL((k_ekte_residual<4>), dim3(g), dim3(256), 0, s, d.n_red, B, scratch, part); break;
    case 8: hipLaunchKernelGGL((k_ekte_residual<8>), dim3(g), dim3(256), 0, s, d.n_red, B, scratch, part); break;
    case 16: hipLaunchKernelGGL((k_ekte_residual<16>), dim3(g), dim3(256), 0, s, d.n_red, B, scratch, part); break;
    default: return CC_E_ARG;
  }
  CC_CHECK(hipGetLastError());
  std::vector<double> hp((size_t)g * 2 * k);
  CC_CHECK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  for (int c = 0; c < k; ++c) {
    double rr = 0.0, bb = 0.0;
    for (int i = 0; i < g; ++i) {
      rr += hp[((size_t)i * 2 + 0) * k + c];
      bb += hp[((size_t)i * 2 + 1) * k + c];
    }
    double bn = std::sqrt(bb);
    if (bn == 0.0) bn = 1.0;
    rel[c] = std::sqrt(rr) / bn;
  }
  return 0;
}

}  // namespace

}  // namespace cc

using namespace cc;

extern "C" {

size_t cc_ekte_workspace_bytes(const cc_ekte_desc* d, int32_t k, int32_t solve) {
  if (!ekte_ok(d) || !k_ok(k)) return 0;
  size_t b = apply_bytes(*d, k);
  if (solve) {
    b += round_up((size_t)256 * 2 * k * sizeof(double)) + 256;  // residual partials
    b += cg_operator_multi_bytes((size_t)d->n_red, k, (size_t)d->n_small) + 512;
  }
  return b;
}

int cc_ekte_apply(const cc_ekte_desc* d, int32_t k, const double* X, double* Y, void* workspace,
                  size_t workspace_bytes, cc_stream_t stream) {
  if (!ekte_ok(d) || !k_ok(k) || !X || !Y || !workspace) return CC_E_ARG;
  if (workspace_bytes < cc_ekte_workspace_bytes(d, k, 0)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const size_t vb = round_up((size_t)(d->n_red + d->n_small) * k * sizeof(double));
  double* XU = (double*)workspace;
  double* YV = (double*)((char*)workspace + vb);
  prof_begin(kProfSpmm, s);
  const int rc = ekte_apply_plain(*d, k, X, Y, XU, YV, s);
  prof_end(s);
  return rc;
}

int cc_pcg_ekte(const cc_ekte_desc* d, int32_t k, const double* B, double* X, double tol, int32_t maxiter,
                int32_t preconditioner, int32_t check_every, void* workspace, size_t workspace_bytes,
                cc_pcg_result* results, cc_stream_t stream) {
  if (!ekte_ok(d) || !k_ok(k) || !B || !X || !workspace || !results || maxiter < 0) return CC_E_ARG;
  if (preconditioner && !d->diag) return CC_E_ARG;
  if (workspace_bytes < cc_ekte_workspace_bytes(d, k, 1)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* base = (char*)workspace;
  const size_t vb = round_up((size_t)(d->n_red + d->n_small) * k * sizeof(double));
  double* XU = (double*)base;
  double* YV = (double*)(base + vb);
  double* rpart = (double*)(base + round_up(apply_bytes(*d, k)));
  char* cgws = (char*)rpart + round_up((size_t)256 * 2 * k * sizeof(double)) + 256;
  const size_t cgb = workspace_bytes - (size_t)(cgws - base);
  const cc_ekte_desc desc = *d;
  const int K = k;
  // P and Q carry n_small tail rows: the gather writes u_s after p, the
  // product writes v_s after q.
  auto apply = [&](const double* P, double* Q, double* partial, int* gs) {
    return ekte_apply(desc, K, const_cast<double*>(P), Q, partial, gs, s);
  };
  auto residual = [&](const double* Xv, double* scratch, double* rel) {
    return ekte_true_residual(desc, K, B, Xv, scratch, XU, YV, rpart, rel, s);
  };
  return cg_operator_multi(desc.n_red, k, apply, preconditioner ? desc.diag : nullptr, B, X, tol, maxiter,
                           preconditioner, check_every, cgws, cgb, results, s, residual, (size_t)desc.n_small);
}

}  // extern "C"
