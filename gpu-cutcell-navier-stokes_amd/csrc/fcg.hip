// fcg.hip - flexible CG with the learned (masked CNN) preconditioner.
//
// Preconditioner: z = D^-1 r + S C^T C S r with S = D^-1/2 and C two masked
// linear 3x3x3 conv layers (cnn.hip): S r is scattered as channel 0 of a
// bfloat16 grid tensor on the background dof grid, the four layers W0, W1,
// flip W1, flip W0 give y = C^T C S r, and z = D^-1 r + S y at the unknowns'
// voxels. In exact arithmetic M is symmetric positive definite with
// M >= D^-1 (Jacobi plus a positive semidefinite learned correction); the
// bfloat16 evaluation makes the applied map slightly inexact, so the outer
// iteration is flexible CG with one stored direction (Notay's FCG(1)):
//   q = A p; alpha = (r.p)/(p.q); x += alpha p; r -= alpha q;
//   z = M(r); beta = -(z.q)/(p.q); p = z + beta p.
// The stopping test, iteration count and breakdown test are those of the
// Jacobi-PCG loop (rel = ||r||/||b|| <= tol; p.q <= 0 or non-finite stops),
// the float64 true residual is recomputed at exit, and the caller falls back
// to Jacobi-PCG when the learned solve does not converge.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {

int cnn_forward_layers(const cc_cnn_desc& net, uint16_t* act0, uint16_t* act1, uint16_t* act2, uint16_t* act3,
                       float* out_f32, hipStream_t s);

struct FcgState {
  double bnorm, rel, alpha, beta, pq, pr, zq;
  int status, iterations;
};

__device__ inline uint16_t f2bf_d(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

template <int NB>
__device__ inline void block_sum_n(double (&v)[NB], double* out, double* red) {
#pragma unroll
  for (int k = 0; k < NB; ++k)
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NB; ++k) red[k * 4 + w] = v[k];
  __syncthreads();
  if (threadIdx.x < NB) out[threadIdx.x] = (red[threadIdx.x * 4] + red[threadIdx.x * 4 + 1]) +
                                           (red[threadIdx.x * 4 + 2] + red[threadIdx.x * 4 + 3]);
}

template <int NB>
__device__ inline void reduce_parts_n(const double* __restrict__ part, int nparts, double* out, double* smem) {
  // part[b][NB]; threads stride over parts, fixed order.
  double acc[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k] += part[(size_t)b * NB + k];
#pragma unroll
  for (int k = 0; k < NB; ++k) smem[k * 256 + threadIdx.x] = acc[k];
  __syncthreads();
  if (threadIdx.x < NB) {
    double s = 0.0;
    for (int i = 0; i < (int)blockDim.x; ++i) s += smem[threadIdx.x * 256 + i];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// Jacobi part D^-1 r (zero diagonal -> identity, as the Jacobi-PCG loop) and
// the symmetric scaling S = D^-1/2 (non-positive diagonal -> 1).
__device__ inline double jacobi_part(double r, double d) { return fabs(d) > 0.0 ? r / d : r; }
__device__ inline double sym_scale(double d) { return d > 0.0 ? 1.0 / sqrt(d) : 1.0; }

// S r written as channel 0 of the grid at the unknowns' voxels.
__global__ __launch_bounds__(256) void k_prec_in(int n, const int* __restrict__ red2full, const double* __restrict__ r,
                                                 const double* __restrict__ diag, uint16_t* __restrict__ grid,
                                                 const FcgState* __restrict__ st) {
  if (st && st->status != CC_RUNNING) return;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    grid[(size_t)red2full[j] * 16] = f2bf_d((float)(sym_scale(diag[j]) * r[j]));
}

// z = D^-1 r + S y (y = channel 0 of the last layer), partial z.q (q may be
// null).
__global__ __launch_bounds__(256) void k_prec_out(int n, const int* __restrict__ red2full, const double* __restrict__ r,
                                                  const double* __restrict__ diag, const float* __restrict__ y,
                                                  double* __restrict__ z, const double* __restrict__ q,
                                                  double* __restrict__ part, const FcgState* __restrict__ st) {
  __shared__ double red[4];
  if (st && st->status != CC_RUNNING) return;
  double zq = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const double d = diag[j];
    const double zz = jacobi_part(r[j], d) + sym_scale(d) * (double)y[(size_t)red2full[j] * 16];
    z[j] = zz;
    if (q) zq += zz * q[j];
  }
  if (part) {
    double v[1] = {zq};
    double out[1];
    block_sum_n<1>(v, out, red);
    if (threadIdx.x == 0) part[blockIdx.x] = out[0];
  }
}

// Setup: x = 0, r = b; partials of b.b.
__global__ __launch_bounds__(256) void k_fcg_setup(int n, const double* __restrict__ b, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ part) {
  __shared__ double red[4];
  double bb = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    x[j] = 0.0;
    r[j] = b[j];
    bb += b[j] * b[j];
  }
  double v[1] = {bb}, out[1];
  block_sum_n<1>(v, out, red);
  if (threadIdx.x == 0) part[blockIdx.x] = out[0];
}

__global__ __launch_bounds__(256) void k_fcg_init(const double* __restrict__ part, int nparts, FcgState* st,
                                                  double* __restrict__ hist) {
  __shared__ double smem[256];
  __shared__ double s[1];
  reduce_parts_n<1>(part, nparts, s, smem);
  if (threadIdx.x == 0) {
    double bn = sqrt(s[0]);
    FcgState f;
    f.bnorm = bn == 0.0 ? 1.0 : bn;
    f.rel = sqrt(s[0]) / f.bnorm;
    f.alpha = f.beta = f.pq = f.pr = f.zq = 0.0;
    f.status = CC_RUNNING;
    f.iterations = 0;
    *st = f;
    if (hist) hist[0] = f.rel;
  }
}

// p = z + beta p (beta = 0 at the start); partials of p.r.
__global__ __launch_bounds__(256) void k_fcg_dir(int n, double* __restrict__ p, const double* __restrict__ z,
                                                 const double* __restrict__ r, double* __restrict__ part,
                                                 const FcgState* __restrict__ st, int first) {
  __shared__ double red[4];
  if (st->status != CC_RUNNING) return;
  const double beta = first ? 0.0 : st->beta;
  double pr = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const double pp = z[j] + beta * (first ? 0.0 : p[j]);
    p[j] = pp;
    pr += pp * r[j];
  }
  double v[1] = {pr}, out[1];
  block_sum_n<1>(v, out, red);
  if (threadIdx.x == 0) part[blockIdx.x] = out[0];
}

// alpha = (p.r)/(p.q) (breakdown when p.q <= 0 or non-finite).
__global__ __launch_bounds__(256) void k_fcg_alpha(const double* __restrict__ part_pq, int n_pq,
                                                   const double* __restrict__ part_pr, int n_pr, FcgState* st,
                                                   int iter) {
  __shared__ double smem[256];
  __shared__ double a[1], b[1];
  if (st->status != CC_RUNNING) return;
  reduce_parts_n<1>(part_pq, n_pq, a, smem);
  reduce_parts_n<1>(part_pr, n_pr, b, smem);
  if (threadIdx.x == 0) {
    const double pq = a[0];
    st->pq = pq;
    st->pr = b[0];
    if (!(pq > 0.0 && isfinite(pq))) {
      st->status = CC_BREAKDOWN;
      st->iterations = iter;
      st->alpha = 0.0;
    } else {
      st->alpha = b[0] / pq;
    }
  }
}

// x += alpha p, r -= alpha q; partials of r.r.
__global__ __launch_bounds__(256) void k_fcg_update(int n, double* __restrict__ x, double* __restrict__ r,
                                                    const double* __restrict__ p, const double* __restrict__ q,
                                                    double* __restrict__ part, const FcgState* __restrict__ st) {
  __shared__ double red[4];
  if (st->status != CC_RUNNING) return;
  const double a = st->alpha;
  double rr = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    x[j] += a * p[j];
    const double rn = r[j] - a * q[j];
    r[j] = rn;
    rr += rn * rn;
  }
  double v[1] = {rr}, out[1];
  block_sum_n<1>(v, out, red);
  if (threadIdx.x == 0) part[blockIdx.x] = out[0];
}

__global__ __launch_bounds__(256) void k_fcg_rel(const double* __restrict__ part, int nparts, FcgState* st, int iter,
                                                 double tol, int maxiter, double* __restrict__ hist) {
  __shared__ double smem[256];
  __shared__ double s[1];
  if (st->status != CC_RUNNING) return;
  reduce_parts_n<1>(part, nparts, s, smem);
  if (threadIdx.x == 0) {
    const double rel = sqrt(s[0]) / st->bnorm;
    st->rel = rel;
    st->iterations = iter;
    if (hist) hist[iter] = rel;
    if (rel <= tol)
      st->status = CC_CONVERGED;
    else if (iter >= maxiter)
      st->status = CC_MAXITER;
  }
}

__global__ __launch_bounds__(256) void k_fcg_beta(const double* __restrict__ part_zq, int n_zq, FcgState* st) {
  __shared__ double smem[256];
  __shared__ double s[1];
  if (st->status != CC_RUNNING) return;
  reduce_parts_n<1>(part_zq, n_zq, s, smem);
  if (threadIdx.x == 0) {
    st->zq = s[0];
    st->beta = -s[0] / st->pq;
  }
}

static int g1(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > (size_t)kMaxParts) g = kMaxParts;
  return (int)(g < 1 ? 1 : g);
}

struct FcgBuffers {
  double *r, *z, *p, *q, *pa, *pb, *pc;
  float* y;
  uint16_t *a0, *a1, *a2, *a3;
  FcgState* st;
};

static size_t fcg_bytes(size_t n, size_t voxels) {
  size_t b = 0;
  auto add = [&](size_t bytes) { b = ((b + 255) & ~size_t(255)) + bytes; };
  for (int i = 0; i < 4; ++i) add(n * sizeof(double));
  for (int i = 0; i < 3; ++i) add((size_t)kMaxParts * sizeof(double));
  add(voxels * 16 * sizeof(float));
  for (int i = 0; i < 4; ++i) add(voxels * 16 * sizeof(uint16_t));
  add(sizeof(FcgState));
  return b + 256;
}

static FcgBuffers carve_fcg(char* base, size_t n, size_t voxels) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = (off + 255) & ~size_t(255);
    char* p = base + off;
    off += bytes;
    return p;
  };
  FcgBuffers b;
  b.r = (double*)take(n * 8);
  b.z = (double*)take(n * 8);
  b.p = (double*)take(n * 8);
  b.q = (double*)take(n * 8);
  b.pa = (double*)take((size_t)kMaxParts * 8);
  b.pb = (double*)take((size_t)kMaxParts * 8);
  b.pc = (double*)take((size_t)kMaxParts * 8);
  b.y = (float*)take(voxels * 16 * 4);
  b.a0 = (uint16_t*)take(voxels * 16 * 2);
  b.a1 = (uint16_t*)take(voxels * 16 * 2);
  b.a2 = (uint16_t*)take(voxels * 16 * 2);
  b.a3 = (uint16_t*)take(voxels * 16 * 2);
  b.st = (FcgState*)take(sizeof(FcgState));
  return b;
}

// The learned preconditioner z = D^-1 r + S C^T C S r; partial z.q.
static int learned_prec(const cc_cnn_desc& net, const FcgBuffers& b, const double* r, const double* diag,
                        const double* q, double* part, hipStream_t s) {
  const int n = net.n_red;
  hipLaunchKernelGGL(k_prec_in, dim3(g1(n)), dim3(256), 0, s, n, net.red2full, r, diag, b.a0, b.st);
  CC_CHECK(cnn_forward_layers(net, b.a0, b.a1, b.a2, b.a3, b.y, s));
  hipLaunchKernelGGL(k_prec_out, dim3(g1(n)), dim3(256), 0, s, n, net.red2full, r, diag, b.y, b.z, q, part, b.st);
  return (int)hipGetLastError();
}

int fcg_operator(int n, const std::function<int(const double*, double*, double*, int*)>& apply,
                 const cc_cnn_desc& net, const double* diag, const double* bvec, double* x, double tol, int maxiter,
                 int check_every, void* ws, size_t wsb, cc_pcg_result* result, double* hist, hipStream_t s,
                 const std::function<int(const double*, double*, double*)>& residual) {
  const size_t voxels = (size_t)net.nx * net.ny * net.nz;
  if (wsb < fcg_bytes(n, voxels)) return CC_E_WORKSPACE;
  FcgBuffers b = carve_fcg((char*)ws, n, voxels);
  // Grid tensors start at zero; only channel 0 at the unknowns' voxels is
  // ever rewritten afterwards.
  CC_CHECK(hipMemsetAsync(b.a0, 0, voxels * 16 * sizeof(uint16_t), s));
  const int gs = g1(n);
  hipLaunchKernelGGL(k_fcg_setup, dim3(gs), dim3(256), 0, s, n, bvec, x, b.r, b.pa);
  hipLaunchKernelGGL(k_fcg_init, dim3(1), dim3(256), 0, s, b.pa, gs, b.st, hist);
  CC_CHECK(learned_prec(net, b, b.r, diag, nullptr, nullptr, s));
  hipLaunchKernelGGL(k_fcg_dir, dim3(gs), dim3(256), 0, s, n, b.p, b.z, b.r, b.pb, b.st, 1);
  CC_CHECK(hipGetLastError());
  FcgState hs{};
  if (check_every < 1) check_every = 1;
  for (int it = 1; it <= maxiter; ++it) {
    int np = 0;
    prof_begin(kProfSpmm, s);
    CC_CHECK(apply(b.p, b.q, b.pa, &np));
    prof_end(s);
    hipLaunchKernelGGL(k_fcg_alpha, dim3(1), dim3(256), 0, s, b.pa, np, b.pb, gs, b.st, it);
    hipLaunchKernelGGL(k_fcg_update, dim3(gs), dim3(256), 0, s, n, x, b.r, b.p, b.q, b.pc, b.st);
    hipLaunchKernelGGL(k_fcg_rel, dim3(1), dim3(256), 0, s, b.pc, gs, b.st, it, tol, maxiter, hist);
    prof_begin(kProfLearned, s);
    CC_CHECK(learned_prec(net, b, b.r, diag, b.q, b.pa, s));
    prof_end(s);
    hipLaunchKernelGGL(k_fcg_beta, dim3(1), dim3(256), 0, s, b.pa, gs, b.st);
    hipLaunchKernelGGL(k_fcg_dir, dim3(gs), dim3(256), 0, s, n, b.p, b.z, b.r, b.pb, b.st, 0);
    CC_CHECK(hipGetLastError());
    if (it % check_every == 0 || it == maxiter) {
      CC_CHECK(hipMemcpyAsync(&hs, b.st, sizeof(FcgState), hipMemcpyDeviceToHost, s));
      CC_CHECK(hipStreamSynchronize(s));
      prof_collect();
      if (hs.status != CC_RUNNING) break;
    }
  }
  CC_CHECK(hipMemcpyAsync(&hs, b.st, sizeof(FcgState), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  prof_collect();
  double rel = 0.0;
  CC_CHECK(residual(x, b.q, &rel));
  result->iterations = hs.iterations;
  result->status = hs.status == CC_RUNNING ? CC_MAXITER : hs.status;
  result->outer = 1;
  result->inner_total = hs.iterations;
  result->recursive_rel = hs.rel;
  result->true_rel = rel;
  return 0;
}

size_t fcg_operator_bytes(size_t n, size_t voxels) { return fcg_bytes(n, voxels); }

// Training input: grid[red2full[j]][0] = bf16(scale * S r_j).
__global__ __launch_bounds__(256) void k_train_input(int n, const int* __restrict__ red2full,
                                                     const float* __restrict__ r, const double* __restrict__ diag,
                                                     float scale, uint16_t* __restrict__ grid) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    grid[(size_t)red2full[j] * 16] = f2bf_d((float)(scale * sym_scale(diag[j]) * (double)r[j]));
}

// Loss ||z - t||^2 / ||t||^2 with z = scale D^-1 r + S y, t = scale e;
// writes dL/dy = 2 (z - t) S / ||t||^2 (times gscale) as channel 0 of the
// bf16 gradient grid and the per-block partials of (||z - t||^2, ||t||^2).
__global__ __launch_bounds__(256) void k_train_loss(int n, const int* __restrict__ red2full,
                                                    const float* __restrict__ r, const float* __restrict__ e,
                                                    const double* __restrict__ diag, float scale,
                                                    const float* __restrict__ y, float inv_den, float gscale,
                                                    uint16_t* __restrict__ gout, double* __restrict__ part) {
  __shared__ double red[8];
  double dd = 0.0, tt = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const double d = diag[j];
    const double sj = sym_scale(d);
    const size_t v = (size_t)red2full[j];
    const double z = scale * jacobi_part((double)r[j], d) + sj * (double)y[v * 16];
    const double t = scale * (double)e[j];
    const double diff = z - t;
    dd += diff * diff;
    tt += t * t;
    gout[v * 16] = f2bf_d((float)(2.0 * diff * sj * inv_den * gscale));
  }
  double v2[2] = {dd, tt}, out[2];
  block_sum_n<2>(v2, out, red);
  if (threadIdx.x < 2) part[2 * blockIdx.x + threadIdx.x] = out[threadIdx.x];
}

}  // namespace cc

using namespace cc;

extern "C" int cc_cnn_train_input(int32_t n, const int32_t* red2full, const float* r, const double* diag, float scale,
                                  uint16_t* grid, cc_stream_t stream) {
  if (n <= 0 || !red2full || !r || !diag || !grid) return CC_E_ARG;
  hipLaunchKernelGGL(k_train_input, dim3(g1(n)), dim3(256), 0, (hipStream_t)stream, n, red2full, r, diag, scale, grid);
  return (int)hipGetLastError();
}

extern "C" int cc_cnn_train_loss(int32_t n, const int32_t* red2full, const float* r, const float* e,
                                 const double* diag, float scale, const float* y, float inv_den, float gscale,
                                 uint16_t* gout, double* partial, int32_t* nparts, cc_stream_t stream) {
  if (n <= 0 || !red2full || !r || !e || !diag || !y || !gout || !partial || !nparts) return CC_E_ARG;
  const int g = g1(n);
  hipLaunchKernelGGL(k_train_loss, dim3(g), dim3(256), 0, (hipStream_t)stream, n, red2full, r, e, diag, scale, y,
                     inv_den, gscale, gout, partial);
  *nparts = g;
  return (int)hipGetLastError();
}

extern "C" int cc_cnn_precondition(const cc_cnn_desc* net, const double* r, const double* diag, double* z,
                                   void* workspace, size_t workspace_bytes, cc_stream_t stream) {
  if (!net || !r || !diag || !z || !workspace || net->n_red <= 0) return CC_E_ARG;
  const size_t voxels = (size_t)net->nx * net->ny * net->nz;
  if (workspace_bytes < fcg_bytes(net->n_red, voxels)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  FcgBuffers b = carve_fcg((char*)workspace, net->n_red, voxels);
  CC_CHECK(hipMemsetAsync(b.a0, 0, voxels * 16 * sizeof(uint16_t), s));
  FcgState init{};
  init.status = CC_RUNNING;
  CC_CHECK(hipMemcpyAsync(b.st, &init, sizeof(FcgState), hipMemcpyHostToDevice, s));
  CC_CHECK(learned_prec(*net, b, r, diag, nullptr, nullptr, s));
  CC_CHECK(hipMemcpyAsync(z, b.z, (size_t)net->n_red * sizeof(double), hipMemcpyDeviceToDevice, s));
  return 0;
}

extern "C" size_t cc_cnn_precondition_workspace_bytes(const cc_cnn_desc* net) {
  if (!net) return 0;
  return fcg_bytes(net->n_red, (size_t)net->nx * net->ny * net->nz);
}
