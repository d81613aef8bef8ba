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
//
// Multilevel form (cc_mlcnn_desc): the background grid and its coarsenings
// (every other grid point per level, mg.hip) each may carry such a network,
// and the correction is summed over the levels through the restriction R and
// the prolongation P = R^T:
//   z = D^-1 r + S sum_l Q_l C_l^T C_l Q_l^T S r,  Q_l = P_1 ... P_l,
// still symmetric and >= D^-1. The coarse levels reach the smooth error
// that a 3x3x3 stack on the fine grid cannot.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {

int cnn_forward_layers(const cc_cnn_desc& net, uint16_t* act0, uint16_t* act1, uint16_t* act2, uint16_t* act3,
                       float* out_f32, hipStream_t s, int nb = 1);
// The layers from a single-channel float32 input to channel 0 of the output
// as a single-channel float32 grid (cnn.hip).
int cnn_forward_ch0(const cc_cnn_desc& net, const float* u, uint16_t* act1, uint16_t* act2, uint16_t* act3, float* y,
                    hipStream_t s, int nb = 1);
// The same four layers on every level with a network, in three launches
// (netpass.hip); net_pass_legacy() selects the per-level layers above.
int net_pass(const cc_mlcnn_desc& ml, float* const* u, uint16_t* const* g1, uint16_t* const* g2, float* const* y,
             int nb, hipStream_t s);
bool net_pass_legacy();

// Grid transfers (mg.hip); nb grids of a batch, each bs elements after the last.
int mg_restrict(const MgDims& g, const float* fine, int fs, float* coarse, int cs, hipStream_t s, int nb = 1,
                size_t fbs = 0, size_t cbs = 0);
int mg_prolong(const MgDims& g, const float* coarse, int cs, const float* add, int as, float* fine, hipStream_t s,
               int nb = 1, size_t cbs = 0, size_t abs_ = 0, size_t fbs = 0);
int mg_to_ch0(long long nvox, const float* src, int ss, uint16_t* grid, hipStream_t s, int nb = 1, size_t sbs = 0,
              size_t gbs = 0);
bool mg_level_dims(const cc_cnn_desc& fine, const cc_cnn_desc& coarse, MgDims* g);

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

// S r written as a single-channel float32 grid at the unknowns' voxels.
__global__ __launch_bounds__(256) void k_prec_in_f32(int n, const int* __restrict__ red2full,
                                                     const double* __restrict__ r, const double* __restrict__ diag,
                                                     float* __restrict__ u0, const FcgState* __restrict__ st) {
  if (st && st->status != CC_RUNNING) return;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    u0[red2full[j]] = (float)(sym_scale(diag[j]) * r[j]);
}

// z = D^-1 r + S y (y[voxel * ys]: channel 0 of the last layer, ys = 16, or
// a single-channel grid, ys = 1), partial z.q (q may be null).
__global__ __launch_bounds__(256) void k_prec_out(int n, const int* __restrict__ red2full, const double* __restrict__ r,
                                                  const double* __restrict__ diag, const float* __restrict__ y,
                                                  int ys, double* __restrict__ z, const double* __restrict__ q,
                                                  double* __restrict__ part, const FcgState* __restrict__ st) {
  __shared__ double red[4];
  if (st && st->status != CC_RUNNING) return;
  double zq = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const double d = diag[j];
    const double zz = jacobi_part(r[j], d) + sym_scale(d) * (double)y[(size_t)red2full[j] * ys];
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
  hipLaunchKernelGGL(k_prec_out, dim3(g1(n)), dim3(256), 0, s, n, net.red2full, r, diag, b.y, 16, b.z, q, part,
                     b.st);
  return (int)hipGetLastError();
}

// The preconditioner of the loop: writes b.z = M(r) and, when q is given,
// the per-block partials of z.q into part.
using PrecFn = std::function<int(const double* r, const double* q, double* part)>;

static int fcg_loop(int n, const std::function<int(const double*, double*, double*, int*)>& apply,
                    const PrecFn& prec, const FcgBuffers& b, const double* bvec, double* x, double tol, int maxiter,
                    int check_every, cc_pcg_result* result, double* hist, hipStream_t s,
                    const std::function<int(const double*, double*, double*)>& residual) {
  const int gs = g1(n);
  hipLaunchKernelGGL(k_fcg_setup, dim3(gs), dim3(256), 0, s, n, bvec, x, b.r, b.pa);
  hipLaunchKernelGGL(k_fcg_init, dim3(1), dim3(256), 0, s, b.pa, gs, b.st, hist);
  CC_CHECK(prec(b.r, nullptr, nullptr));
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
    CC_CHECK(prec(b.r, b.q, b.pa));
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
  PrecFn prec = [&](const double* r, const double* q, double* part) {
    return learned_prec(net, b, r, diag, q, part, s);
  };
  return fcg_loop(n, apply, prec, b, bvec, x, tol, maxiter, check_every, result, hist, s, residual);
}

// ---------------------------------------------------------------------------
// Multilevel learned preconditioner.
// ---------------------------------------------------------------------------

struct MlBuffers {
  float* u[CC_ML_MAX_LEVELS];  // S r restricted to level l (single channel)
  float* t[CC_ML_MAX_LEVELS];  // the correction prolongated down to level l
  float* y[CC_ML_MAX_LEVELS];  // channel 0 of the network output (levels with a net)
  uint16_t* act[CC_ML_MAX_LEVELS][4];  // [1..3]: hidden activations [voxel][16]
};

static size_t ml_vox(const cc_cnn_desc& d) { return (size_t)d.nx * d.ny * d.nz; }

bool ml_desc_ok(const cc_mlcnn_desc& ml) {
  if (ml.levels < 1 || ml.levels > CC_ML_MAX_LEVELS) return false;
  if (ml.net[0].n_red <= 0 || !ml.net[0].red2full) return false;
  if (!ml.has_net[ml.levels - 1]) return false;  // the coarsest level carries a network
  for (int l = 0; l < ml.levels; ++l) {
    const cc_cnn_desc& d = ml.net[l];
    if (d.nx <= 0 || d.ny <= 0 || d.nz <= 0) return false;
    if (ml.has_net[l] && (!d.mask || !d.weights[0] || !d.weights[1] || !d.weights[2] || !d.weights[3])) return false;
    if (l > 0) {
      MgDims g;
      if (!mg_level_dims(ml.net[l - 1], d, &g)) return false;
      for (int k = 0; k < 3; ++k)
        if (d.periodic[k] != ml.net[l - 1].periodic[k]) return false;
    }
  }
  return true;
}

size_t ml_bytes(const cc_mlcnn_desc& ml) {
  size_t b = 0;
  auto add = [&](size_t bytes) { b = ((b + 255) & ~size_t(255)) + bytes; };
  for (int l = 0; l < ml.levels; ++l) {
    const size_t v = ml_vox(ml.net[l]);
    add(v * sizeof(float));
    add(v * sizeof(float));
    if (ml.has_net[l]) {
      add(v * sizeof(float));
      for (int k = 1; k < 4; ++k) add(v * 16 * sizeof(uint16_t));
    }
  }
  return b + 256;
}

static MlBuffers carve_ml(const cc_mlcnn_desc& ml, char* base) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = (off + 255) & ~size_t(255);
    char* p = base + off;
    off += bytes;
    return p;
  };
  MlBuffers b{};
  for (int l = 0; l < ml.levels; ++l) {
    const size_t v = ml_vox(ml.net[l]);
    b.u[l] = (float*)take(v * sizeof(float));
    b.t[l] = (float*)take(v * sizeof(float));
    if (ml.has_net[l]) {
      b.y[l] = (float*)take(v * sizeof(float));
      for (int k = 1; k < 4; ++k) b.act[l][k] = (uint16_t*)take(v * 16 * sizeof(uint16_t));
    }
  }
  return b;
}

// Zero what is written only in part: u[0] off the unknowns' voxels.
static int ml_setup(const cc_mlcnn_desc& ml, const MlBuffers& b, hipStream_t s) {
  CC_CHECK(hipMemsetAsync(b.u[0], 0, ml_vox(ml.net[0]) * sizeof(float), s));
  return 0;
}

// y[l] = C_l^T C_l u[l] on every level with a network, nb grids per level.
static int level_networks(const cc_mlcnn_desc& ml, const MlBuffers& mb, int nb, hipStream_t s) {
  if (!net_pass_legacy()) {
    uint16_t* g1[CC_ML_MAX_LEVELS] = {};
    uint16_t* g2[CC_ML_MAX_LEVELS] = {};
    for (int l = 0; l < ml.levels; ++l) {
      g1[l] = mb.act[l][1];
      g2[l] = mb.act[l][2];
    }
    return net_pass(ml, mb.u, g1, g2, mb.y, nb, s);
  }
  for (int l = 0; l < ml.levels; ++l) {
    if (!ml.has_net[l]) continue;
    CC_CHECK(cnn_forward_ch0(ml.net[l], mb.u[l], mb.act[l][1], mb.act[l][2], mb.act[l][3], mb.y[l], s, nb));
  }
  return 0;
}

// z = D^-1 r + S sum_l Q_l C_l^T C_l Q_l^T S r; partial z.q when q is given.
// (Running the coarse chain on a second stream beside the fine level's
// network was measured slower: the fine convolutions fill the GPU, and the
// fork/join events added about 100 us per application.)
static int ml_prec(const cc_mlcnn_desc& ml, const MlBuffers& mb, const double* r, const double* diag,
                   const double* q, double* z, double* part, const FcgState* st, hipStream_t s) {
  const cc_cnn_desc& f = ml.net[0];
  const int n = f.n_red;
  const int L = ml.levels;
  hipLaunchKernelGGL(k_prec_in_f32, dim3(g1(n)), dim3(256), 0, s, n, f.red2full, r, diag, mb.u[0], st);
  MgDims g[CC_ML_MAX_LEVELS];
  for (int l = 1; l < L; ++l) {
    mg_level_dims(ml.net[l - 1], ml.net[l], &g[l]);
    CC_CHECK(mg_restrict(g[l], mb.u[l - 1], 1, mb.u[l], 1, s));
  }
  CC_CHECK(level_networks(ml, mb, 1, s));
  const float* src = mb.y[L - 1];
  for (int l = L - 1; l >= 1; --l) {
    CC_CHECK(mg_prolong(g[l], src, 1, ml.has_net[l - 1] ? mb.y[l - 1] : nullptr, 1, mb.t[l - 1], s));
    src = mb.t[l - 1];
  }
  hipLaunchKernelGGL(k_prec_out, dim3(g1(n)), dim3(256), 0, s, n, f.red2full, r, diag, src, 1, z, q, part, st);
  return (int)hipGetLastError();
}

int fcg_operator_ml(int n, const std::function<int(const double*, double*, double*, int*)>& apply,
                    const cc_mlcnn_desc& ml, const double* diag, const double* bvec, double* x, double tol,
                    int maxiter, int check_every, void* ws, size_t wsb, cc_pcg_result* result, double* hist,
                    hipStream_t s, const std::function<int(const double*, double*, double*)>& residual) {
  const size_t core = (fcg_bytes(n, 0) + 255) & ~size_t(255);
  if (wsb < core + ml_bytes(ml)) return CC_E_WORKSPACE;
  FcgBuffers b = carve_fcg((char*)ws, n, 0);
  MlBuffers mb = carve_ml(ml, (char*)ws + core);
  CC_CHECK(ml_setup(ml, mb, s));
  PrecFn prec = [&](const double* r, const double* q, double* part) {
    return ml_prec(ml, mb, r, diag, q, b.z, part, b.st, s);
  };
  return fcg_loop(n, apply, prec, b, bvec, x, tol, maxiter, check_every, result, hist, s, residual);
}

size_t fcg_operator_ml_bytes(size_t n, const cc_mlcnn_desc& ml) {
  return ((fcg_bytes(n, 0) + 255) & ~size_t(255)) + ml_bytes(ml);
}

size_t fcg_operator_bytes(size_t n, size_t voxels) { return fcg_bytes(n, voxels); }

// ---------------------------------------------------------------------------
// Flexible CG on K right-hand sides at once (row-major [n][K] vectors, one
// FcgState per column), with the multilevel learned preconditioner applied
// to all K columns in the same launches: the grid transfers and the
// convolutions run over K grids per launch (batched), so the operator's
// K-column apply and the network's weights are shared by the batch. Each
// column follows its own FCG(1) recurrence and stops on its own; stopped
// columns are left untouched.
// ---------------------------------------------------------------------------

// Sum v over the threads of a 256-thread block that share column
// threadIdx.x % K; thread c < K writes the block's total to out[c].
template <int K>
__device__ inline void block_cols(double v, double* out, double* red) {
#pragma unroll
  for (int o = K; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane < K) red[w * K + lane] = v;
  __syncthreads();
  if (threadIdx.x < K)
    out[threadIdx.x] = (red[threadIdx.x] + red[K + threadIdx.x]) + (red[2 * K + threadIdx.x] + red[3 * K + threadIdx.x]);
}

// Per-column fixed-order sum of part[nparts][K] by one block of
// kFinK = 1024 threads: each thread's loads are all issued before they are
// summed (unrolled), then a fixed-order tree over the threads of a column.
constexpr int kFinK = 1024;
template <int K>
__device__ inline void reduce_cols(const double* __restrict__ part, int nparts, double* out, double* smem) {
  const int c = threadIdx.x % K, r0 = threadIdx.x / K;
  constexpr int R = kFinK / K;
  double acc = 0.0;
#pragma unroll 8
  for (int b = r0; b < nparts; b += R) acc += part[(size_t)b * K + c];
  smem[threadIdx.x] = acc;
  __syncthreads();
  for (int h = R / 2; h > 0; h >>= 1) {
    if (r0 < h) smem[threadIdx.x] += smem[threadIdx.x + h * K];
    __syncthreads();
  }
  if (threadIdx.x < K) out[threadIdx.x] = smem[threadIdx.x];
  __syncthreads();
}

template <int K>
__global__ __launch_bounds__(256) void k_fcgk_setup(int n, const double* __restrict__ b, double* __restrict__ x,
                                                    double* __restrict__ r, double* __restrict__ part) {
  __shared__ double red[4 * K];
  double bb = 0.0;
  const size_t total = (size_t)n * K;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    x[e] = 0.0;
    r[e] = b[e];
    bb += b[e] * b[e];
  }
  block_cols<K>(bb, part + (size_t)blockIdx.x * K, red);
}

// A zero right-hand side is solved by x = 0 (converged at iteration 0).
template <int K>
__global__ __launch_bounds__(kFinK) void k_fcgk_init(const double* __restrict__ part, int nparts, FcgState* st) {
  __shared__ double smem[kFinK];
  __shared__ double s[K];
  reduce_cols<K>(part, nparts, s, smem);
  if (threadIdx.x < K) {
    const double bb = s[threadIdx.x];
    FcgState f;
    f.bnorm = bb > 0.0 ? sqrt(bb) : 1.0;
    f.rel = sqrt(bb) / f.bnorm;
    f.alpha = f.beta = f.pq = f.pr = f.zq = 0.0;
    f.status = bb > 0.0 ? CC_RUNNING : CC_CONVERGED;
    f.iterations = 0;
    st[threadIdx.x] = f;
  }
}

// p = z + beta p (beta = 0 at the start) on running columns; partials p.r.
template <int K>
__global__ __launch_bounds__(256) void k_fcgk_dir(int n, double* __restrict__ p, const double* __restrict__ z,
                                                  const double* __restrict__ r, double* __restrict__ part,
                                                  const FcgState* __restrict__ st, int first) {
  __shared__ double red[4 * K];
  const int c = threadIdx.x % K;
  const bool run = st[c].status == CC_RUNNING;
  const double beta = first ? 0.0 : st[c].beta;
  double pr = 0.0;
  const size_t total = (size_t)n * K;
  if (run)
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
      const double pp = z[e] + (first ? 0.0 : beta * p[e]);
      p[e] = pp;
      pr += pp * r[e];
    }
  block_cols<K>(pr, part + (size_t)blockIdx.x * K, red);
}

template <int K>
__global__ __launch_bounds__(kFinK) void k_fcgk_alpha(const double* __restrict__ part_pq, int n_pq,
                                                    const double* __restrict__ part_pr, int n_pr, FcgState* st,
                                                    int iter) {
  __shared__ double smem[kFinK];
  __shared__ double a[K], b[K];
  reduce_cols<K>(part_pq, n_pq, a, smem);
  reduce_cols<K>(part_pr, n_pr, b, smem);
  if (threadIdx.x < K) {
    FcgState& f = st[threadIdx.x];
    if (f.status == CC_RUNNING) {
      const double pq = a[threadIdx.x];
      f.pq = pq;
      f.pr = b[threadIdx.x];
      if (!(pq > 0.0 && isfinite(pq))) {
        f.status = CC_BREAKDOWN;
        f.iterations = iter;
        f.alpha = 0.0;
      } else {
        f.alpha = b[threadIdx.x] / pq;
      }
    }
  }
}

template <int K>
__global__ __launch_bounds__(kFinK) void k_fcgk_rel(const double* __restrict__ part, int nparts, FcgState* st, int iter,
                                                  double tol, int maxiter) {
  __shared__ double smem[kFinK];
  __shared__ double s[K];
  reduce_cols<K>(part, nparts, s, smem);
  if (threadIdx.x < K) {
    FcgState& f = st[threadIdx.x];
    if (f.status == CC_RUNNING) {
      const double rel = sqrt(s[threadIdx.x]) / f.bnorm;
      f.rel = rel;
      f.iterations = iter;
      if (rel <= tol)
        f.status = CC_CONVERGED;
      else if (iter >= maxiter)
        f.status = CC_MAXITER;
    }
  }
}

template <int K>
__global__ __launch_bounds__(kFinK) void k_fcgk_beta(const double* __restrict__ part_zq, int n_zq, FcgState* st) {
  __shared__ double smem[kFinK];
  __shared__ double s[K];
  reduce_cols<K>(part_zq, n_zq, s, smem);
  if (threadIdx.x < K) {
    FcgState& f = st[threadIdx.x];
    if (f.status == CC_RUNNING) {
      f.zq = s[threadIdx.x];
      f.beta = -s[threadIdx.x] / f.pq;
    }
  }
}

// S r of every column as single-channel float32 grids u0[c][voxel].
template <int K>
__global__ __launch_bounds__(256) void k_preck_in(int n, const int* __restrict__ red2full, const double* __restrict__ r,
                                                  const double* __restrict__ diag, float* __restrict__ u0,
                                                  size_t vox) {
  // Through an LDS tile of 1024 / K unknowns x K columns: the reads of r are
  // whole K-column rows, the writes runs of one column's grid.
  constexpr int TJ = 1024 / K;
  __shared__ float tile[K][TJ + 1];
  const size_t ntiles = ((size_t)n + TJ - 1) / TJ;
  for (size_t tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
    const size_t j0 = tb * TJ;
    __syncthreads();  // the previous tile's runs are written
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = threadIdx.x + 256 * q;
      const int ji = k / K, c = k % K;
      const size_t j = j0 + ji;
      if (j < (size_t)n) tile[c][ji] = (float)(sym_scale(diag[j]) * r[j * K + c]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = threadIdx.x + 256 * q;
      const int c = k / TJ, ji = k % TJ;
      const size_t j = j0 + ji;
      if (j < (size_t)n) u0[(size_t)c * vox + red2full[j]] = tile[c][ji];
    }
  }
}

// The flexible-CG update fused with the scaling into the grids: on running
// columns x += alpha p and r -= alpha q (partials r.r per column, for the
// stop test), then S r of every column into the grids u0[c][voxel], as
// k_preck_in scales them, with one read of r. At most kMaxParts blocks
// (grid-strided over the tiles).
template <int K>
__global__ __launch_bounds__(256) void k_fcgk_update_in(int n, const int* __restrict__ red2full,
                                                        double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ p, const double* __restrict__ q,
                                                        const FcgState* __restrict__ st,
                                                        const double* __restrict__ diag, float* __restrict__ u0,
                                                        size_t vox, double* __restrict__ part) {
  constexpr int TJ = 1024 / K;
  __shared__ float tile[K][TJ + 1];
  __shared__ double red[4 * K];
  const int c = threadIdx.x % K;  // the column of every element this thread updates
  const bool run = st[c].status == CC_RUNNING;
  const double a = st[c].alpha;
  double rr = 0.0;
  const size_t ntiles = ((size_t)n + TJ - 1) / TJ;
  for (size_t tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
    const size_t j0 = tb * TJ;
    __syncthreads();  // the previous tile's runs are written
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int k = threadIdx.x + 256 * qq;
      const int ji = k / K;
      const size_t j = j0 + ji;
      if (j < (size_t)n) {
        const size_t e = j * K + c;
        double rn = r[e];
        if (run) {
          x[e] += a * p[e];
          rn = rn - a * q[e];
          r[e] = rn;
          rr += rn * rn;
        }
        tile[c][ji] = (float)(sym_scale(diag[j]) * rn);
      }
    }
    __syncthreads();
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int k = threadIdx.x + 256 * qq;
      const int cc = k / TJ, ji = k % TJ;
      const size_t j = j0 + ji;
      if (j < (size_t)n) u0[(size_t)cc * vox + red2full[j]] = tile[cc][ji];
    }
  }
  block_cols<K>(rr, part + (size_t)blockIdx.x * K, red);
}

// z = D^-1 r + S y0; partials z.q per column. Without `coarse`, y0 is
// column c's fine grid y at c * ybs (element stride ys). With `coarse`, the
// last prolongation is done here: y0 = y + P coarse_c (y may be null),
// coarse_c the next level's grid at c * cbs and P that of g, summed as
// k_mg_prolong sums it, so y0 never goes through memory.
template <int K>
__global__ __launch_bounds__(256) void k_preck_out(int n, const int* __restrict__ red2full, const double* __restrict__ r,
                                                   const double* __restrict__ diag, const float* __restrict__ y,
                                                   int ys, size_t ybs, const float* __restrict__ coarse, MgDims g,
                                                   size_t cbs, double* __restrict__ z, const double* __restrict__ q,
                                                   double* __restrict__ part) {
  // Through an LDS tile of 1024 / K unknowns x K columns: the reads of y are
  // runs of one column's grid, the rows of r, q and z whole K-column rows.
  constexpr int TJ = 1024 / K;
  __shared__ float tile[TJ][K + 1];
  __shared__ double red[4 * K];
  // Per unknown of the tile, the prolongation's taps (shared by the K
  // columns): coarse offsets and weights, tap-major so that a wave's lanes
  // (consecutive unknowns) read consecutive words.
  __shared__ int s_off[8][TJ];
  __shared__ float s_w[8][TJ];
  __shared__ int s_v[TJ];
  const int c = threadIdx.x % K;  // the column of every element this thread writes
  const size_t ntiles = ((size_t)n + TJ - 1) / TJ;
  double zq = 0.0;
  for (size_t tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
    const size_t j0 = tb * TJ;
    __syncthreads();  // the previous tile's rows are done
    if (coarse) {
      if (threadIdx.x < TJ && j0 + threadIdx.x < (size_t)n) {
        const int v = red2full[j0 + threadIdx.x];
        const int nfx = g.a[0].nf, nfy = g.a[1].nf;
        const int fx = v % nfx, row = v / nfx, fy = row % nfy, fz = row / nfy;
        size_t off[8];
        float w[8];
        prolong_taps(g, fx, fy, fz, off, w);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          s_off[t][threadIdx.x] = (int)off[t];
          s_w[t][threadIdx.x] = w[t];
        }
        s_v[threadIdx.x] = v;
      }
      __syncthreads();
    }
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int k = threadIdx.x + 256 * qq;
      const int cc = k / TJ, ji = k % TJ;
      const size_t j = j0 + ji;
      if (j < (size_t)n) {
        if (coarse) {
          // As prolong_value sums: the add term, then the 8 taps in order.
          const float* cb = coarse + (size_t)cc * cbs;
          float acc = y ? y[(size_t)cc * ybs + (size_t)s_v[ji] * ys] : 0.0f;
          float cv[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) cv[t] = cb[s_off[t][ji]];
#pragma unroll
          for (int t = 0; t < 8; ++t) acc += s_w[t][ji] * cv[t];
          tile[ji][cc] = acc;
        } else {
          tile[ji][cc] = y[(size_t)cc * ybs + (size_t)red2full[j] * ys];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
      const int ji = (threadIdx.x + 256 * qq) / K;
      const size_t j = j0 + ji;
      if (j < (size_t)n) {
        const size_t e = j * K + c;
        const double d = diag[j];
        const double zz = jacobi_part(r[e], d) + sym_scale(d) * (double)tile[ji][c];
        z[e] = zz;
        if (q) zq += zz * q[e];
      }
    }
  }
  if (part) block_cols<K>(zq, part + (size_t)blockIdx.x * K, red);
}

static int gk(size_t n, int K) {
  size_t g = (n * K + 1023) / 1024;
  if (g > (size_t)kMaxParts) g = kMaxParts;
  return (int)(g < 1 ? 1 : g);
}

struct FcgKBuffers {
  double *r, *z, *p, *q, *pa, *pb, *pc;
  FcgState* st;
};

static size_t fcgk_bytes(size_t n, int K) {
  size_t b = 0;
  auto add = [&](size_t bytes) { b = ((b + 255) & ~size_t(255)) + bytes; };
  for (int i = 0; i < 4; ++i) add(n * K * sizeof(double));
  for (int i = 0; i < 3; ++i) add((size_t)kMaxParts * K * sizeof(double));
  add(K * sizeof(FcgState));
  return b + 256;
}

static FcgKBuffers carve_fcgk(char* base, size_t n, int K) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = (off + 255) & ~size_t(255);
    char* p = base + off;
    off += bytes;
    return p;
  };
  FcgKBuffers b;
  b.r = (double*)take(n * K * 8);
  b.z = (double*)take(n * K * 8);
  b.p = (double*)take(n * K * 8);
  b.q = (double*)take(n * K * 8);
  b.pa = (double*)take((size_t)kMaxParts * K * 8);
  b.pb = (double*)take((size_t)kMaxParts * K * 8);
  b.pc = (double*)take((size_t)kMaxParts * K * 8);
  b.st = (FcgState*)take(K * sizeof(FcgState));
  return b;
}

// Multilevel buffers for K columns: every per-level array of MlBuffers
// holds K grids one after the other.
static size_t ml_bytes_k(const cc_mlcnn_desc& ml, int K) {
  size_t b = 0;
  auto add = [&](size_t bytes) { b = ((b + 255) & ~size_t(255)) + bytes; };
  for (int l = 0; l < ml.levels; ++l) {
    const size_t v = ml_vox(ml.net[l]) * K;
    add(v * sizeof(float));
    add(v * sizeof(float));
    if (ml.has_net[l]) {
      add(v * sizeof(float));
      for (int k = 1; k < 4; ++k) add(v * 16 * sizeof(uint16_t));
    }
  }
  return b + 256;
}

static MlBuffers carve_ml_k(const cc_mlcnn_desc& ml, char* base, int K) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = (off + 255) & ~size_t(255);
    char* p = base + off;
    off += bytes;
    return p;
  };
  MlBuffers b{};
  for (int l = 0; l < ml.levels; ++l) {
    const size_t v = ml_vox(ml.net[l]) * K;
    b.u[l] = (float*)take(v * sizeof(float));
    b.t[l] = (float*)take(v * sizeof(float));
    if (ml.has_net[l]) {
      b.y[l] = (float*)take(v * sizeof(float));
      for (int k = 1; k < 4; ++k) b.act[l][k] = (uint16_t*)take(v * 16 * sizeof(uint16_t));
    }
  }
  return b;
}

// scaled_in: the grids u0 already hold S r (k_fcgk_update_in wrote them).
template <int K>
static int ml_prec_k(const cc_mlcnn_desc& ml, const MlBuffers& mb, const double* r, const double* diag,
                     const double* q, double* z, double* part, hipStream_t s, bool scaled_in = false) {
  const cc_cnn_desc& f = ml.net[0];
  const int n = f.n_red;
  const int L = ml.levels;
  size_t vox[CC_ML_MAX_LEVELS];
  for (int l = 0; l < L; ++l) vox[l] = ml_vox(ml.net[l]);
  // One block per 1024-element tile (no partials).
  if (!scaled_in)
    hipLaunchKernelGGL((k_preck_in<K>), dim3((unsigned)(((size_t)n * K + 1023) / 1024)), dim3(256), 0, s, n,
                       f.red2full, r, diag, mb.u[0], vox[0]);
  MgDims g[CC_ML_MAX_LEVELS] = {};
  for (int l = 1; l < L; ++l) {
    mg_level_dims(ml.net[l - 1], ml.net[l], &g[l]);
    CC_CHECK(mg_restrict(g[l], mb.u[l - 1], 1, mb.u[l], 1, s, K, vox[l - 1], vox[l]));
  }
  CC_CHECK(level_networks(ml, mb, K, s));
  const float* src = mb.y[L - 1];
  for (int l = L - 1; l >= 2; --l) {
    CC_CHECK(mg_prolong(g[l], src, 1, ml.has_net[l - 1] ? mb.y[l - 1] : nullptr, 1, mb.t[l - 1], s, K, vox[l],
                        vox[l - 1], vox[l - 1]));
    src = mb.t[l - 1];
  }
  // The prolongation to the finest grid is done by the output kernel, at
  // the active voxels only.
  if (L >= 2)
    hipLaunchKernelGGL((k_preck_out<K>), dim3(gk(n, K)), dim3(256), 0, s, n, f.red2full, r, diag,
                       ml.has_net[0] ? mb.y[0] : nullptr, 1, vox[0], src, g[1], vox[1], z, q, part);
  else
    hipLaunchKernelGGL((k_preck_out<K>), dim3(gk(n, K)), dim3(256), 0, s, n, f.red2full, r, diag, src, 1, vox[0],
                       nullptr, g[0], (size_t)0, z, q, part);
  return (int)hipGetLastError();
}

template <int K>
static int fcg_loop_ml_k(int n, const std::function<int(const double*, double*, double*, int*)>& apply,
                         const cc_mlcnn_desc& ml, const MlBuffers& mb, const FcgKBuffers& b, const double* diag,
                         const double* bvec, double* x, double tol, int maxiter, int check_every,
                         cc_pcg_result* results, hipStream_t s,
                         const std::function<int(const double*, double*, double*)>& residual) {
  const int gs = gk(n, K);
  hipLaunchKernelGGL((k_fcgk_setup<K>), dim3(gs), dim3(256), 0, s, n, bvec, x, b.r, b.pa);
  hipLaunchKernelGGL((k_fcgk_init<K>), dim3(1), dim3(kFinK), 0, s, b.pa, gs, b.st);
  CC_CHECK(ml_prec_k<K>(ml, mb, b.r, diag, nullptr, b.z, nullptr, s));
  hipLaunchKernelGGL((k_fcgk_dir<K>), dim3(gs), dim3(256), 0, s, n, b.p, b.z, b.r, b.pb, b.st, 1);
  CC_CHECK(hipGetLastError());
  std::vector<FcgState> hs(K);
  if (check_every < 1) check_every = 1;
  for (int it = 1; it <= maxiter; ++it) {
    int np = 0;
    prof_begin(kProfSpmm, s);
    CC_CHECK(apply(b.p, b.q, b.pa, &np));
    prof_end(s);
    hipLaunchKernelGGL((k_fcgk_alpha<K>), dim3(1), dim3(kFinK), 0, s, b.pa, np, b.pb, gs, b.st, it);
    // The update also scales r into the finest grids (the preconditioner's
    // first step), so the timed preconditioner starts at the restrictions.
    hipLaunchKernelGGL((k_fcgk_update_in<K>), dim3(gs), dim3(256), 0, s, n, ml.net[0].red2full, x, b.r, b.p, b.q,
                       b.st, diag, mb.u[0], ml_vox(ml.net[0]), b.pc);
    hipLaunchKernelGGL((k_fcgk_rel<K>), dim3(1), dim3(kFinK), 0, s, b.pc, gs, b.st, it, tol, maxiter);
    prof_begin(kProfLearned, s);
    CC_CHECK(ml_prec_k<K>(ml, mb, b.r, diag, b.q, b.z, b.pa, s, true));
    prof_end(s);
    hipLaunchKernelGGL((k_fcgk_beta<K>), dim3(1), dim3(kFinK), 0, s, b.pa, gs, b.st);
    hipLaunchKernelGGL((k_fcgk_dir<K>), dim3(gs), dim3(256), 0, s, n, b.p, b.z, b.r, b.pb, b.st, 0);
    CC_CHECK(hipGetLastError());
    if (it % check_every == 0 || it == maxiter) {
      CC_CHECK(hipMemcpyAsync(hs.data(), b.st, K * sizeof(FcgState), hipMemcpyDeviceToHost, s));
      CC_CHECK(hipStreamSynchronize(s));
      prof_collect();
      bool any = false;
      for (int c = 0; c < K; ++c) any |= hs[c].status == CC_RUNNING;
      if (!any) break;
    }
  }
  CC_CHECK(hipMemcpyAsync(hs.data(), b.st, K * sizeof(FcgState), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  prof_collect();
  std::vector<double> rel(K, 0.0);
  CC_CHECK(residual(x, b.q, rel.data()));
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

size_t fcg_operator_ml_k_bytes(size_t n, const cc_mlcnn_desc& ml, int K) {
  return ((fcgk_bytes(n, K) + 255) & ~size_t(255)) + ml_bytes_k(ml, K);
}

int fcg_operator_ml_k(int n, int K, const std::function<int(const double*, double*, double*, int*)>& apply,
                      const cc_mlcnn_desc& ml, const double* diag, const double* bvec, double* x, double tol,
                      int maxiter, int check_every, void* ws, size_t wsb, cc_pcg_result* results, hipStream_t s,
                      const std::function<int(const double*, double*, double*)>& residual) {
  if (wsb < fcg_operator_ml_k_bytes(n, ml, K)) return CC_E_WORKSPACE;
  const size_t core = (fcgk_bytes(n, K) + 255) & ~size_t(255);
  FcgKBuffers b = carve_fcgk((char*)ws, n, K);
  MlBuffers mb = carve_ml_k(ml, (char*)ws + core, K);
  // Zero what is written only in part: u[0] off the unknowns' voxels, for
  // all K grids.
  CC_CHECK(hipMemsetAsync(mb.u[0], 0, ml_vox(ml.net[0]) * K * sizeof(float), s));
  switch (K) {
    case 1: return fcg_loop_ml_k<1>(n, apply, ml, mb, b, diag, bvec, x, tol, maxiter, check_every, results, s, residual);
    case 2: return fcg_loop_ml_k<2>(n, apply, ml, mb, b, diag, bvec, x, tol, maxiter, check_every, results, s, residual);
    case 4: return fcg_loop_ml_k<4>(n, apply, ml, mb, b, diag, bvec, x, tol, maxiter, check_every, results, s, residual);
    case 8: return fcg_loop_ml_k<8>(n, apply, ml, mb, b, diag, bvec, x, tol, maxiter, check_every, results, s, residual);
    case 16:
      return fcg_loop_ml_k<16>(n, apply, ml, mb, b, diag, bvec, x, tol, maxiter, check_every, results, s, residual);
    default: return CC_E_ARG;
  }
}

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

// Multilevel training input: u0[red2full[j]] = scale * S r_j (float32).
__global__ __launch_bounds__(256) void k_mg_train_input(int n, const int* __restrict__ red2full,
                                                        const float* __restrict__ r, const double* __restrict__ diag,
                                                        float scale, float* __restrict__ u0) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    u0[red2full[j]] = (float)(scale * sym_scale(diag[j]) * (double)r[j]);
}

// Multilevel loss: z = scale D^-1 r + S t0 (t0 the summed correction on the
// background grid, single channel), L = ||z - scale e||^2 / ||scale e||^2;
// g0 = dL/dt0 (times gscale) at the unknowns' voxels; partials as k_train_loss.
__global__ __launch_bounds__(256) void k_mg_train_loss(int n, const int* __restrict__ red2full,
                                                       const float* __restrict__ r, const float* __restrict__ e,
                                                       const double* __restrict__ diag, float scale,
                                                       const float* __restrict__ t0, float inv_den, float gscale,
                                                       float* __restrict__ g0, double* __restrict__ part) {
  __shared__ double red[8];
  double dd = 0.0, tt = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const double d = diag[j];
    const double sj = sym_scale(d);
    const size_t v = (size_t)red2full[j];
    const double z = scale * jacobi_part((double)r[j], d) + sj * (double)t0[v];
    const double t = scale * (double)e[j];
    const double diff = z - t;
    dd += diff * diff;
    tt += t * t;
    g0[v] = (float)(2.0 * diff * sj * inv_den * gscale);
  }
  double v2[2] = {dd, tt}, out[2];
  block_sum_n<2>(v2, out, red);
  if (threadIdx.x < 2) part[2 * blockIdx.x + threadIdx.x] = out[threadIdx.x];
}

}  // namespace cc

using namespace cc;

extern "C" int cc_mg_train_input(int32_t n, const int32_t* red2full, const float* r, const double* diag, float scale,
                                 float* u0, cc_stream_t stream) {
  if (n <= 0 || !red2full || !r || !diag || !u0) return CC_E_ARG;
  hipLaunchKernelGGL(k_mg_train_input, dim3(g1(n)), dim3(256), 0, (hipStream_t)stream, n, red2full, r, diag, scale,
                     u0);
  return (int)hipGetLastError();
}

extern "C" int cc_mg_train_loss(int32_t n, const int32_t* red2full, const float* r, const float* e,
                                const double* diag, float scale, const float* t0, float inv_den, float gscale,
                                float* g0, double* partial, int32_t* nparts, cc_stream_t stream) {
  if (n <= 0 || !red2full || !r || !e || !diag || !t0 || !g0 || !partial || !nparts) return CC_E_ARG;
  const int g = g1(n);
  hipLaunchKernelGGL(k_mg_train_loss, dim3(g), dim3(256), 0, (hipStream_t)stream, n, red2full, r, e, diag, scale, t0,
                     inv_den, gscale, g0, partial);
  *nparts = g;
  return (int)hipGetLastError();
}

extern "C" size_t cc_mlcnn_precondition_workspace_bytes(const cc_mlcnn_desc* ml) {
  if (!ml || !ml_desc_ok(*ml)) return 0;
  return ((fcg_bytes(ml->net[0].n_red, 0) + 255) & ~size_t(255)) + ml_bytes(*ml);
}

extern "C" int cc_mlcnn_precondition(const cc_mlcnn_desc* ml, const double* r, const double* diag, double* z,
                                     void* workspace, size_t workspace_bytes, cc_stream_t stream) {
  if (!ml || !ml_desc_ok(*ml) || !r || !diag || !z || !workspace) return CC_E_ARG;
  if (workspace_bytes < cc_mlcnn_precondition_workspace_bytes(ml)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const int n = ml->net[0].n_red;
  const size_t core = (fcg_bytes(n, 0) + 255) & ~size_t(255);
  FcgBuffers b = carve_fcg((char*)workspace, n, 0);
  MlBuffers mb = carve_ml(*ml, (char*)workspace + core);
  CC_CHECK(ml_setup(*ml, mb, s));
  FcgState init{};
  init.status = CC_RUNNING;
  CC_CHECK(hipMemcpyAsync(b.st, &init, sizeof(FcgState), hipMemcpyHostToDevice, s));
  return ml_prec(*ml, mb, r, diag, nullptr, z, nullptr, b.st, s);
}

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
