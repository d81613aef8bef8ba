// gmres.hip - restarted GMRES with the Jacobi preconditioner on a CSR matrix
// (float64, one right-hand side): the GMRES leg of the solver interface.
//
// It follows SciPy 1.15.3's gmres as the reference's solve_cpu_gmres calls it
// (code/fast_solver_backends.py:135-191; scipy/sparse/linalg/_isolve/
// iterative.py, gmres): left preconditioning (v0 = M r / ||M r||,
// w = M A v), modified Gram-Schmidt, Givens rotations in LAPACK's dlartg
// convention, the inner stop on the preconditioned residual estimate
// presid <= ptol with ptol adapted between restarts (gh-8400 tolerance
// control), the outer stop on the true residual ||b - A x|| <= tol ||b||,
// maxiter counting restarts, and the iteration count being the number of
// inner steps (the pr_norm callbacks). The Jacobi preconditioner is the CPU
// rule of code/fast_solver_backends.py:74-87 (zero diagonal -> 0).
//
// The device does the n-long work: the product (the autotuned CSR SpMV), the
// preconditioner, every Gram-Schmidt step and the x update. A Gram-Schmidt
// step is one launch: each block re-reduces the previous dot's per-block
// partials in a fixed order (so every block holds the same h), applies
// w -= h v_k to its rows and writes its partial of the next dot. The
// (restart + 1) x restart Hessenberg problem stays on the host, which reads
// one column per inner step.

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {

namespace {

constexpr int kGmBlock = 256;
constexpr int kGmMaxBlocks = 512;

int gm_grid(int n) {
  int g = (n + kGmBlock - 1) / kGmBlock;
  return g < 1 ? 1 : (g > kGmMaxBlocks ? kGmMaxBlocks : g);
}

__device__ inline double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  }
  __syncthreads();
  return s;
}

// Fixed-order sum of nparts partials, the same in every block.
__device__ inline double sum_parts(const double* __restrict__ part, int nparts, double* red) {
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) v += part[i];
  v = block_sum(v, red);
  __shared__ double bcast;
  if (threadIdx.x == 0) bcast = v;
  __syncthreads();
  return bcast;
}

// dst = inv .* src (inv[i] = 1 / d_i, or 0 for a zero diagonal; all ones
// without a preconditioner); partials of ||dst||^2 and, when v0 is set,
// of v0 . dst.
__global__ __launch_bounds__(kGmBlock) void k_gm_psolve(int n, const double* __restrict__ inv,
                                                        const double* __restrict__ src, double* __restrict__ dst,
                                                        const double* __restrict__ v0, double* __restrict__ part_sq,
                                                        double* __restrict__ part_dot) {
  __shared__ double red[kGmBlock / 64];
  double sq = 0.0, dt = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double w = inv ? inv[i] * src[i] : src[i];
    dst[i] = w;
    sq += w * w;
    if (v0) dt += v0[i] * w;
  }
  sq = block_sum(sq, red);
  if (threadIdx.x == 0) part_sq[blockIdx.x] = sq;
  if (v0) {
    dt = block_sum(dt, red);
    if (threadIdx.x == 0) part_dot[blockIdx.x] = dt;
  }
}

// One Gram-Schmidt step: h = sum(part_in) (= v_k . w), w -= h v_k, then the
// partials of v_next . w (v_next set) or of ||w||^2 (v_next null) into
// part_out; block 0 writes h to *h_out.
__global__ __launch_bounds__(kGmBlock) void k_gm_mgs(int n, const double* __restrict__ vk,
                                                     const double* __restrict__ part_in, int nparts,
                                                     double* __restrict__ w, const double* __restrict__ v_next,
                                                     double* __restrict__ part_out, double* __restrict__ h_out) {
  __shared__ double red[kGmBlock / 64];
  const double h = sum_parts(part_in, nparts, red);
  if (blockIdx.x == 0 && threadIdx.x == 0) *h_out = h;
  double acc = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double wi = w[i] - h * vk[i];
    w[i] = wi;
    acc += v_next ? v_next[i] * wi : wi * wi;
  }
  acc = block_sum(acc, red);
  if (threadIdx.x == 0) part_out[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kGmBlock) void k_gm_scale(int n, double s, const double* __restrict__ src,
                                                       double* __restrict__ dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i] * s;
}

// x += sum_k y[k] v_k (k ascending).
__global__ __launch_bounds__(kGmBlock) void k_gm_update(int n, int m, const double* __restrict__ y,
                                                        const double* __restrict__ V, size_t ldv,
                                                        double* __restrict__ x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < m; ++k) s += y[k] * V[(size_t)k * ldv + i];
    x[i] += s;
  }
}

// r = b - Ax; partials of ||r||^2.
__global__ __launch_bounds__(kGmBlock) void k_gm_resid(int n, const double* __restrict__ b,
                                                       const double* __restrict__ ax, double* __restrict__ r,
                                                       double* __restrict__ part) {
  __shared__ double red[kGmBlock / 64];
  double sq = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double v = b[i] - ax[i];
    r[i] = v;
    sq += v * v;
  }
  sq = block_sum(sq, red);
  if (threadIdx.x == 0) part[blockIdx.x] = sq;
}

// inv[i] = 1 / A_ii where A_ii != 0, else 0 (the CPU Jacobi rule).
__global__ __launch_bounds__(kGmBlock) void k_gm_inv_diag(int n, const int* __restrict__ indptr,
                                                          const int* __restrict__ indices,
                                                          const double* __restrict__ vals, double* __restrict__ inv) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double d = 0.0;
    for (int e = indptr[i]; e < indptr[i + 1]; ++e)
      if (indices[e] == i) d += vals[e];
    inv[i] = d != 0.0 ? 1.0 / d : 0.0;
  }
}

// LAPACK dlartg (3.10+): c, s, r with [c s; -s c] [f; g] = [r; 0].
void lartg(double f, double g, double& c, double& s, double& r) {
  const double safmin = DBL_MIN;
  const double safmax = 1.0 / safmin;
  const double rtmin = std::sqrt(safmin);
  const double rtmax = std::sqrt(safmax / 2.0);
  const double f1 = std::fabs(f), g1 = std::fabs(g);
  if (g == 0.0) {
    c = 1.0;
    s = 0.0;
    r = f;
  } else if (f == 0.0) {
    c = 0.0;
    s = std::copysign(1.0, g);
    r = g1;
  } else if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
    const double d = std::sqrt(f * f + g * g);
    c = f1 / d;
    r = std::copysign(d, f);
    s = g / r;
  } else {
    const double u = std::fmin(safmax, std::fmax(safmin, std::fmax(f1, g1)));
    const double fs = f / u, gs = g / u;
    const double d = std::sqrt(fs * fs + gs * gs);
    c = std::fabs(fs) / d;
    r = std::copysign(d, f);
    s = gs / r;
    r *= u;
  }
}

struct GmBuffers {
  double *V, *w, *r, *ax, *inv, *pa, *pb, *pc, *hcol, *y;
  size_t ldv;
};

size_t gm_bytes(int n, int restart) {
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t ld = up((size_t)n * sizeof(double)) / sizeof(double);
  size_t b = up((size_t)(restart + 1) * ld * sizeof(double));
  b += 4 * up((size_t)n * sizeof(double));                  // w, r, ax, inv
  b += 3 * up((size_t)kGmMaxBlocks * sizeof(double));       // partials
  b += 2 * up((size_t)(restart + 2) * sizeof(double));      // H column, y
  return b + 256;
}

GmBuffers gm_carve(char* base, int n, int restart) {
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  GmBuffers g;
  g.ldv = up((size_t)n * sizeof(double)) / sizeof(double);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    double* p = reinterpret_cast<double*>(base + off);
    off += up(bytes);
    return p;
  };
  g.V = take((size_t)(restart + 1) * g.ldv * sizeof(double));
  g.w = take((size_t)n * sizeof(double));
  g.r = take((size_t)n * sizeof(double));
  g.ax = take((size_t)n * sizeof(double));
  g.inv = take((size_t)n * sizeof(double));
  g.pa = take((size_t)kGmMaxBlocks * sizeof(double));
  g.pb = take((size_t)kGmMaxBlocks * sizeof(double));
  g.pc = take((size_t)kGmMaxBlocks * sizeof(double));
  g.hcol = take((size_t)(restart + 2) * sizeof(double));
  g.y = take((size_t)(restart + 2) * sizeof(double));
  return g;
}

double host_sum(const double* dev, int count, hipStream_t s, int* err) {
  std::vector<double> h(count);
  *err = (int)hipMemcpyAsync(h.data(), dev, count * sizeof(double), hipMemcpyDeviceToHost, s);
  if (!*err) *err = (int)hipStreamSynchronize(s);
  double t = 0.0;
  for (double v : h) t += v;
  return t;
}

}  // namespace

int gmres_run(int n, const int* indptr, const int* indices, const double* vals, const double* b, double* x,
              double tol, int restart, int maxiter, int precond, void* ws, cc_gmres_result* out, hipStream_t s) {
  GmBuffers g = gm_carve((char*)ws, n, restart);
  const int G = gm_grid(n);
  int err = 0;
  // The mean row length picks the default SpMV shape (when none is tuned).
  int ends[2] = {0, 0};
  CC_CHECK(hipMemcpyAsync(&ends[0], indptr, sizeof(int), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipMemcpyAsync(&ends[1], indptr + n, sizeof(int), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  const double avg = double(ends[1] - ends[0]) / n;
  const double* inv = nullptr;
  if (precond) {
    hipLaunchKernelGGL(k_gm_inv_diag, dim3(G), dim3(kGmBlock), 0, s, n, indptr, indices, vals, g.inv);
    inv = g.inv;
  }
  // ||b|| and ||M b||.
  hipLaunchKernelGGL(k_gm_psolve, dim3(G), dim3(kGmBlock), 0, s, n, nullptr, b, g.r, nullptr, g.pa, nullptr);
  const double bnrm2 = std::sqrt(host_sum(g.pa, G, s, &err));
  CC_CHECK(err);
  std::memset(out, 0, sizeof(*out));
  out->b_norm = bnrm2;
  hipLaunchKernelGGL(k_gm_scale, dim3(G), dim3(kGmBlock), 0, s, n, 0.0, b, x);  // x0 = 0
  if (bnrm2 == 0.0) {
    out->residual_norm = 0.0;
    return (int)hipStreamSynchronize(s);
  }
  const double atol = tol * bnrm2;
  if (bnrm2 < atol) {  // tol > 1: x0 = 0 already meets it
    out->residual_norm = bnrm2;
    return (int)hipStreamSynchronize(s);
  }
  const double eps = DBL_EPSILON;
  if (maxiter <= 0) maxiter = 10 * n;
  if (restart <= 0) restart = 20;
  if (restart > n) restart = n;
  hipLaunchKernelGGL(k_gm_psolve, dim3(G), dim3(kGmBlock), 0, s, n, inv, b, g.w, nullptr, g.pa, nullptr);
  const double Mb_nrm2 = std::sqrt(host_sum(g.pa, G, s, &err));
  CC_CHECK(err);
  double ptol_max_factor = 1.0;
  double ptol = Mb_nrm2 * std::fmin(ptol_max_factor, atol / bnrm2);
  double presid = 0.0, rnorm = bnrm2;
  std::vector<double> h((size_t)restart * (restart + 1), 0.0);  // h[col][k]
  std::vector<double> gc(restart), gsn(restart), S(restart + 1), y(restart + 1);
  auto H = [&](int col, int k) -> double& { return h[(size_t)col * (restart + 1) + k]; };
  int inner = 0, outer = 0;
  // r = b (x0 = 0).
  CC_CHECK(hipMemcpyAsync(g.r, b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
  for (int iteration = 0; iteration < maxiter; ++iteration) {
    ++outer;
    // v0 = M r / ||M r||.
    hipLaunchKernelGGL(k_gm_psolve, dim3(G), dim3(kGmBlock), 0, s, n, inv, g.r, g.V, nullptr, g.pa, nullptr);
    double tmp = std::sqrt(host_sum(g.pa, G, s, &err));
    CC_CHECK(err);
    hipLaunchKernelGGL(k_gm_scale, dim3(G), dim3(kGmBlock), 0, s, n, 1.0 / tmp, g.V, g.V);
    std::fill(S.begin(), S.end(), 0.0);
    S[0] = tmp;
    bool breakdown = false;
    int col = 0;
    for (col = 0; col < restart; ++col) {
      double* vcol = g.V + (size_t)col * g.ldv;
      CC_CHECK(csr_spmm_f64(n, 1, avg, indptr, indices, vals, vcol, g.ax, nullptr, s, nullptr));
      // w = M A v; partials of ||w||^2 (h0) and of v_0 . w.
      hipLaunchKernelGGL(k_gm_psolve, dim3(G), dim3(kGmBlock), 0, s, n, inv, g.ax, g.w, g.V, g.pc, g.pa);
      double* pin = g.pa;
      double* pout = g.pb;
      for (int k = 0; k <= col; ++k) {
        const double* vnext = k < col ? g.V + (size_t)(k + 1) * g.ldv : nullptr;
        hipLaunchKernelGGL(k_gm_mgs, dim3(G), dim3(kGmBlock), 0, s, n, g.V + (size_t)k * g.ldv, pin, G, g.w, vnext,
                           pout, g.hcol + k);
        std::swap(pin, pout);
      }
      CC_CHECK(hipGetLastError());
      // Host: the column's dots, ||w|| before (h0) and after (h1) the sweep.
      std::vector<double> hc(col + 1), p1(G), p0(G);
      CC_CHECK(hipMemcpyAsync(hc.data(), g.hcol, (col + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
      CC_CHECK(hipMemcpyAsync(p1.data(), pin, G * sizeof(double), hipMemcpyDeviceToHost, s));
      CC_CHECK(hipMemcpyAsync(p0.data(), g.pc, G * sizeof(double), hipMemcpyDeviceToHost, s));
      CC_CHECK(hipStreamSynchronize(s));
      double h0 = 0.0, h1 = 0.0;
      for (int i = 0; i < G; ++i) {
        h0 += p0[i];
        h1 += p1[i];
      }
      h0 = std::sqrt(h0);
      h1 = std::sqrt(h1);
      for (int k = 0; k <= col; ++k) H(col, k) = hc[k];
      H(col, col + 1) = h1;
      double* vnew = g.V + (size_t)(col + 1) * g.ldv;
      if (h1 <= eps * h0) {
        H(col, col + 1) = 0.0;
        breakdown = true;
        CC_CHECK(hipMemcpyAsync(vnew, g.w, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s));
      } else {
        hipLaunchKernelGGL(k_gm_scale, dim3(G), dim3(kGmBlock), 0, s, n, 1.0 / h1, g.w, vnew);
      }
      for (int k = 0; k < col; ++k) {
        const double c = gc[k], sn = gsn[k];
        const double n0 = H(col, k), n1 = H(col, k + 1);
        H(col, k) = c * n0 + sn * n1;
        H(col, k + 1) = -sn * n0 + c * n1;
      }
      double c, sn, mag;
      lartg(H(col, col), H(col, col + 1), c, sn, mag);
      gc[col] = c;
      gsn[col] = sn;
      H(col, col) = mag;
      H(col, col + 1) = 0.0;
      const double t2 = -sn * S[col];
      S[col] = c * S[col];
      S[col + 1] = t2;
      presid = std::fabs(t2);
      ++inner;
      if (presid <= ptol || breakdown) break;
    }
    if (col == restart) --col;
    // Upper triangular solve of the Hessenberg problem (pseudo-solve of
    // singular cases as SciPy does).
    if (H(col, col) == 0.0) S[col] = 0.0;
    for (int k = 0; k <= col; ++k) y[k] = S[k];
    for (int k = col; k > 0; --k) {
      if (y[k] != 0.0) {
        y[k] /= H(k, k);
        const double t = y[k];
        for (int j = 0; j < k; ++j) y[j] -= t * H(k, j);
      }
    }
    if (y[0] != 0.0) y[0] /= H(0, 0);
    CC_CHECK(hipMemcpyAsync(g.y, y.data(), (col + 1) * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gm_update, dim3(G), dim3(kGmBlock), 0, s, n, col + 1, g.y, g.V, g.ldv, x);
    CC_CHECK(csr_spmm_f64(n, 1, avg, indptr, indices, vals, x, g.ax, nullptr, s, nullptr));
    hipLaunchKernelGGL(k_gm_resid, dim3(G), dim3(kGmBlock), 0, s, n, b, g.ax, g.r, g.pa);
    rnorm = std::sqrt(host_sum(g.pa, G, s, &err));
    CC_CHECK(err);
    if (rnorm <= atol) break;
    if (breakdown) break;
    if (presid <= ptol)
      ptol_max_factor = std::fmax(eps, 0.25 * ptol_max_factor);
    else
      ptol_max_factor = std::fmin(1.0, 1.5 * ptol_max_factor);
    ptol = presid * std::fmin(ptol_max_factor, atol / rnorm);
  }
  out->iterations = inner;
  out->restarts = outer;
  out->info = rnorm <= atol ? 0 : maxiter;
  out->residual_norm = rnorm;
  return 0;
}

size_t gmres_bytes(int n, int restart) { return gm_bytes(n, restart); }

}  // namespace cc

using namespace cc;

extern "C" {

size_t cc_gmres_workspace_bytes(int32_t n, int32_t restart) {
  if (n <= 0) return 0;
  if (restart <= 0) restart = 20;
  if (restart > n) restart = n;
  return gmres_bytes(n, restart);
}

int cc_gmres_jacobi_f64(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                        const double* b, double* x, double tol, int32_t restart, int32_t maxiter,
                        int32_t preconditioner, void* workspace, size_t workspace_bytes, cc_gmres_result* result,
                        cc_stream_t stream) {
  if (n <= 0 || !indptr || !indices || !vals || !b || !x || !workspace || !result || maxiter < 0 || restart < 0 ||
      !(tol >= 0.0))
    return CC_E_ARG;
  int rs = restart <= 0 ? 20 : restart;
  if (rs > n) rs = n;
  if (workspace_bytes < gmres_bytes(n, rs)) return CC_E_WORKSPACE;
  return gmres_run(n, indptr, indices, vals, b, x, tol, rs, maxiter, preconditioner, workspace, result,
                   (hipStream_t)stream);
}

}  // extern "C"
