// sumfact.hip - matrix-free aggregated operator y = E^T (K + c M) E x.
//
// Launch sequence of one apply (all on the caller's stream):
//   G  u = E x on the background dof grid (interior copy, small-dof stencils)
//   T  v = A_all u: the tensor-product operator of every background cell,
//      Kz(x)My(x)Mx + Mz(x)Ky(x)Mx + Mz(x)My(x)(Kx + c Mx), by sum factorisation.
//      A workgroup owns an 8x8x8 output tile, stages the (8+2p)^3 input halo in
//      LDS and runs three banded 1D passes (x: a = Mx u, b = (Kx + c Mx) u;
//      y: c = My b + Ky a, d = My a; z: v = Mz c + Kz d), each from LDS.
//   S  per shell cell s (cut, or empty but touching an active dof), one wave:
//      W_s = C_s u_s - A_s u_s (cut) or -A_s u_s (empty), A_s the cell's
//      full tensor block built from its 1D element forms.
//   C  v[g] += sum of the W entries of the shell cells containing g (CSR over
//      touched dofs, fixed order: deterministic, no atomics).
//   E  y = E^T v, with per-block partials of x.y for the CG loop.
// The operator is the same as the explicit E^T A E of the reference's
// restrict_spd (code/bell_3d_core.py:550-552) but never formed.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {

// Correctly rounded multiply-add and product in the value type T of an apply
// (float64 for the solver, float32 for the inner solve of the mixed one).
__device__ inline double fma_t(double a, double b, double c) { return __fma_rn(a, b, c); }
__device__ inline float fma_t(float a, float b, float c) { return __fmaf_rn(a, b, c); }
__device__ inline double mul_t(double a, double b) { return __dmul_rn(a, b); }
__device__ inline float mul_t(float a, float b) { return __fmul_rn(a, b); }

__device__ inline int wrap_or_invalid(int i, int n, int periodic) {
  if (periodic) {
    i %= n;
    return i < 0 ? i + n : i;
  }
  return (i >= 0 && i < n) ? i : -1;
}

// u = E x, interior dofs: u[g] = x[full2red[g]] (0 off the unknowns; the
// small dofs are written by k_mf_gather_small afterwards).
__global__ __launch_bounds__(256) void k_mf_gather(cc_sumfact_desc d, const double* __restrict__ x,
                                                   double* __restrict__ u) {
  const int n_full = d.nb[0] * d.nb[1] * d.nb[2];
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n_full; g += gridDim.x * blockDim.x) {
    const int r = d.full2red[g];
    u[g] = r >= 0 ? x[r] : 0.0;
  }
}

// u = E x, small dofs: one wave per small dof, its stencil entries spread
// over the lanes (a stencil holds up to a few tens of entries) and summed by
// a butterfly, so a small dof costs two memory latencies, not one per entry.
__global__ __launch_bounds__(256) void k_mf_gather_small(cc_sumfact_desc d, const double* __restrict__ x,
                                                         double* __restrict__ u) {
  const int lane = threadIdx.x & 63;
  const int sm = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (sm >= d.n_small) return;
  double s = 0.0;
  for (int t = d.small_ptr[sm] + lane; t < d.small_ptr[sm + 1]; t += 64) s += d.small_w[t] * x[d.small_col[t]];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) u[d.small_dof[sm]] = s;
}

// Tensor-product part on the background grid, v = (Kx My Mz + Mx Ky Mz +
// Mx My Kz + c Mx My Mz) u with the banded 1D forms (x pass a = Mx u,
// b = (Kx + c Mx) u; y pass c = My b + Ky a, d = My a; z pass v = Mz c +
// Kz d), streamed along z: a block owns a 16 x 16 (x, y) column of the grid
// and a run of zc output planes. For each input
// plane it stages the plane's 22 x 22 halo in LDS (loaded into registers
// one plane ahead), runs the x and y passes for that plane, and keeps the
// y-pass results of the last 2p + 1 planes in an LDS ring, from which the z
// pass emits output plane z - p. Every input value is staged once per
// (x, y) tile and chunk instead of once per 8^3 tile with a (8 + 2p)^3 halo.
// blockIdx.z = column * nzc + chunk (column c of u_all / v_all starts
// c * n_full in).
constexpr int kSx = 16, kSy = 16;
constexpr int kMaxZc = 32;  // output planes per block at most

template <typename T, int P1>
__global__ __launch_bounds__(256) void k_mf_tensor_z(cc_sumfact_desc d, const T* __restrict__ u_all,
                                                     T* __restrict__ v_all, int nzc, int zc) {
  constexpr int p = P1 - 1;
  constexpr int W = 2 * p + 1;
  constexpr int HX = kSx + 2 * p, HY = kSy + 2 * p;
  constexpr int HALO = HX * HY;
  constexpr int PER = (HALO + 255) / 256;
  __shared__ T su[HY][HX];
  __shared__ T sa[HY][kSx];
  __shared__ T sb[HY][kSx];
  // The y-pass results of the last W planes: each thread reads back only its
  // own (x, y), so the ring lives in registers, indexed by plane mod W (the
  // plane loop is unrolled by W so that every index is a constant).
  T rc[W], rd[W];
  const int nbx = d.nb[0], nby = d.nb[1], nbz = d.nb[2];
  const size_t plane = (size_t)nbx * nby;
  // Blocks go to the 8 XCDs round-robin by linear id; each XCD is given one
  // contiguous range of the (x, y, z-run, column) tile space instead, so
  // that adjacent columns, whose halos overlap, share that XCD's L2.
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned total = gx * gy * gridDim.z;
  const unsigned lin = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  constexpr unsigned kX = 8;  // XCDs of the part
  const unsigned xq = total / kX, xr = total % kX, xcd = lin % kX;
  const unsigned tile = xcd * xq + min(xcd, xr) + lin / kX;
  const int bx = (int)(tile % gx), by = (int)((tile / gx) % gy), bz = (int)(tile / (gx * gy));
  const size_t col_off = (size_t)(bz / nzc) * plane * nbz;
  const T* __restrict__ u = u_all + col_off;
  T* __restrict__ v = v_all + col_off;
  const int zs = (bz % nzc) * zc;
  const int ze = min(nbz, zs + zc);
  const int x0 = bx * kSx, y0 = by * kSy;
  const int tid = threadIdx.x;
  const int tx = tid % kSx, ty = tid / kSx;
  const double cm = d.mass_coef;
  // This thread's halo elements: in-plane offset, or -1 outside the grid.
  long long hoff[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = tid + 256 * k;
    hoff[k] = -1;
    if (e < HALO) {
      const int gx = wrap_or_invalid(x0 - p + e % HX, nbx, d.periodic[0]);
      const int gy = wrap_or_invalid(y0 - p + e / HX, nby, d.periodic[1]);
      if (gx >= 0 && gy >= 0) hoff[k] = (long long)gy * nbx + gx;
    }
  }
  auto fetch = [&](int zi, T (&val)[PER]) {
    const int gz = wrap_or_invalid(zi, nbz, d.periodic[2]);
#pragma unroll
    for (int k = 0; k < PER; ++k) val[k] = (gz >= 0 && hoff[k] >= 0) ? u[(size_t)gz * plane + hoff[k]] : T(0);
  };
  const int gxo = x0 + tx, gyo = y0 + ty;
  const int z_first = zs - p;
  // Band rows held for the whole z run: this thread's x row (both of its
  // x-pass items have x = x0 + tx) and y row in registers, the z rows of the
  // block's output planes in LDS.
  T mx[W], kx[W], my[W], ky[W];
#pragma unroll
  for (int k = 0; k < W; ++k) {
    mx[k] = gxo < nbx ? (T)d.band_m[0][gxo * W + k] : T(0);
    kx[k] = gxo < nbx ? (T)(d.band_k[0][gxo * W + k] + cm * d.band_m[0][gxo * W + k]) : T(0);
    my[k] = gyo < nby ? (T)d.band_m[1][gyo * W + k] : T(0);
    ky[k] = gyo < nby ? (T)d.band_k[1][gyo * W + k] : T(0);
  }
  __shared__ T mz[kMaxZc][W], kz[kMaxZc][W];
  for (int i = tid; i < (ze - zs) * W; i += 256) {
    mz[i / W][i % W] = (T)d.band_m[2][(size_t)zs * W + i];
    kz[i / W][i % W] = (T)d.band_k[2][(size_t)zs * W + i];
  }
  T val[PER];
  fetch(z_first, val);
  for (int zb = z_first; zb < ze + p; zb += W) {
#pragma unroll
    for (int u = 0; u < W; ++u) {
      const int zi = zb + u;
      if (zi >= ze + p) break;
      // Every thread has finished the previous plane's x pass (the barrier
      // after it) and its y pass (before the first barrier below), so su,
      // and then sa and sb, may be rewritten.
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = tid + 256 * k;
        if (e < HALO) su[e / HX][e % HX] = val[k];
      }
      if (zi + 1 < ze + p) fetch(zi + 1, val);
      __syncthreads();
      // x pass: a = Mx u, b = (Kx + c Mx) u on the HY rows of the halo.
      for (int i = tid; i < HY * kSx; i += 256) {
        const int ly = i / kSx, lx = i % kSx;  // lx == tx
        T a = 0, b = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) {
          const T uu = su[ly][lx + k];
          a += mx[k] * uu;
          b += kx[k] * uu;
        }
        sa[ly][lx] = a;
        sb[ly][lx] = b;
      }
      __syncthreads();
      // y pass into the ring slot of this plane ((zi - z_first) mod W = u):
      // c = My b + Ky a, d = My a.
      {
        T cc_ = 0, dd = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) {
          cc_ += my[k] * sb[ty + k][tx] + ky[k] * sa[ty + k][tx];
          dd += my[k] * sa[ty + k][tx];
        }
        rc[u] = cc_;
        rd[u] = dd;
      }
      // z pass: output plane zo = zi - p from ring planes zo - p .. zo + p,
      // slot (u - 2p + k) mod W for k = 0 .. 2p.
      const int zo = zi - p;
      if (zo >= zs && gxo < nbx && gyo < nby) {
        T out = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) {
          constexpr int kW = W;
          const int sl = ((u - 2 * p + k) % kW + kW) % kW;
          out += mz[zo - zs][k] * rc[sl] + kz[zo - zs][k] * rd[sl];
        }
        v[(size_t)zo * plane + (size_t)gyo * nbx + gxo] = out;
      }
    }
  }
}

// z-chunk length: enough blocks (about 2048) with the shortest chunk
// (4 planes, 2p + 4 input planes each) for small grids.
static int tensor_zc(const cc_sumfact_desc& d, int K) {
  const long long xy = (long long)((d.nb[0] + kSx - 1) / kSx) * ((d.nb[1] + kSy - 1) / kSy);
  long long zc = (long long)d.nb[2] * xy * K / 2048;
  if (zc < 4) zc = 4;
  if (zc > kMaxZc) zc = kMaxZc;
  return (int)zc;
}

template <typename T, int P1>
static void tensor_launch(const cc_sumfact_desc& d, const T* u, T* v, int K, hipStream_t s) {
  const int zc = tensor_zc(d, K);
  const int nzc = (d.nb[2] + zc - 1) / zc;
  const dim3 grid((d.nb[0] + kSx - 1) / kSx, (d.nb[1] + kSy - 1) / kSy, nzc * K);
  hipLaunchKernelGGL((k_mf_tensor_z<T, P1>), grid, dim3(256), 0, s, d, u, v, nzc, zc);
}

// v[g] += sum of the shell-cell outputs W that belong to dof g; a 16-lane
// team per touched dof (a dof lies in up to (p+1)^3 shell cells).
__global__ __launch_bounds__(256) void k_mf_touch(cc_sumfact_desc d, const double* __restrict__ Win,
                                                  double* __restrict__ v) {
  const int tl = threadIdx.x & 15;
  const int team = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int nteams = (gridDim.x * blockDim.x) >> 4;
  for (int t = team; t < d.n_touch; t += nteams) {
    double s = 0.0;
    for (int i = d.touch_ptr[t] + tl; i < d.touch_ptr[t + 1]; i += 16) s += Win[d.touch_idx[i]];
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    if (tl == 0) v[d.touch_dof[t]] += s;
  }
}

// The E^T part of the small dofs: extra[j] = sum_b E[b, j] v[small dof b]
// for the unknowns j that have such entries (16-lane team per unknown).
__global__ __launch_bounds__(256) void k_mf_et_extra(cc_sumfact_desc d, const double* __restrict__ v,
                                                     double* __restrict__ extra, const ColState* __restrict__ st) {
  if (st && st[0].status != CC_RUNNING) return;
  const int tl = threadIdx.x & 15;
  const int team = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int nteams = (gridDim.x * blockDim.x) >> 4;
  for (int i = team; i < d.n_et_rows; i += nteams) {
    const int j = d.et_rows[i];
    double s = 0.0;
    for (int t = d.et_ptr[j] + tl; t < d.et_ptr[j + 1]; t += 16) s += d.et_w[t] * v[d.small_dof[d.et_small[t]]];
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    if (tl == 0) extra[j] = s;
  }
}

// y = E^T v (interior part plus extra[] where the unknown has E^T entries);
// partial[block] = sum_j x_j y_j (when partial != null).
__global__ __launch_bounds__(256) void k_mf_scatter(cc_sumfact_desc d, const double* __restrict__ v,
                                                    const double* __restrict__ extra, double* __restrict__ y,
                                                    const double* __restrict__ x, double* __restrict__ partial,
                                                    const ColState* __restrict__ st) {
  __shared__ double red[4];
  if (st && st[0].status != CC_RUNNING) return;
  double dot = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < d.n_red; j += gridDim.x * blockDim.x) {
    double s = v[d.red2full[j]];
    if (d.et_ptr[j + 1] > d.et_ptr[j]) s += extra[j];
    y[j] = s;
    if (partial) dot += x[j] * s;
  }
  if (partial) {
    for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dot;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
  }
}

// Residual r = b - y and partials of (r.r, b.b) per block.
__global__ __launch_bounds__(256) void k_mf_residual(int n, const double* __restrict__ b, const double* __restrict__ y,
                                                     double* __restrict__ partial) {
  __shared__ double red[2][4];
  double rr = 0.0, bb = 0.0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const double r = b[j] - y[j];
    rr += r * r;
    bb += b[j] * b[j];
  }
  for (int o = 32; o > 0; o >>= 1) {
    rr += __shfl_xor(rr, o, 64);
    bb += __shfl_xor(bb, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = rr;
    red[1][threadIdx.x >> 6] = bb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    partial[2 * blockIdx.x + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
  }
}

// ---------------------------------------------------------------------------
// K right-hand sides. Reduced-space blocks are row-major [n_red][K] (as the
// CSR solver's); the background-grid work arrays u, v are column-major
// [K][n_full], so that each column's tensor apply reads and writes
// contiguous memory, and every shell cell reads its cut block once for all K
// columns. Shell outputs are [n_corr][K][P], extra is [n_red][K].
// ---------------------------------------------------------------------------

// u[c][g] = x[full2red[g]][c] (0 off the interior), through an LDS tile of
// 1024 / K dofs x K columns: the reads are whole K-column rows of x, the
// writes runs of consecutive dofs of one column.
template <typename T, int K>
__global__ __launch_bounds__(256) void k_mfk_gather(cc_sumfact_desc d, const T* __restrict__ x,
                                                    T* __restrict__ u) {
  constexpr int TG = 1024 / K;
  __shared__ T tile[K][TG + 1];
  const size_t n_full = (size_t)d.nb[0] * d.nb[1] * d.nb[2];
  const size_t ntiles = (n_full + TG - 1) / TG;
  for (size_t tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
    const size_t g0 = tb * TG;
    __syncthreads();  // the previous tile's column runs are written
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = threadIdx.x + 256 * q;
      const int gi = k / K, c = k % K;
      const size_t g = g0 + gi;
      T val = 0;
      if (g < n_full) {
        const int r = d.full2red[g];
        if (r >= 0) val = x[(size_t)r * K + c];
      }
      tile[c][gi] = val;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = threadIdx.x + 256 * q;
      const int c = k / TG, gi = k % TG;
      const size_t g = g0 + gi;
      if (g < n_full) u[(size_t)c * n_full + g] = tile[c][gi];
    }
  }
}

// One wave per small dof: lane (entry slot e, column c), 64 / K entry slots.
template <typename T, int K>
__global__ __launch_bounds__(256) void k_mfk_gather_small(cc_sumfact_desc d, const T* __restrict__ x,
                                                          T* __restrict__ u) {
  constexpr int EL = 64 / K;
  const int lane = threadIdx.x & 63;
  const int c = lane % K, el = lane / K;
  const int sm = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (sm >= d.n_small) return;
  const size_t n_full = (size_t)d.nb[0] * d.nb[1] * d.nb[2];
  T s = 0;
  const int tend = d.small_ptr[sm + 1];
  int t = d.small_ptr[sm] + el;
  // Four entries' loads in flight at once, summed in entry order.
  for (; t + 3 * EL < tend; t += 4 * EL) {
    int col[4];
    T w[4], xv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) col[q] = d.small_col[t + q * EL];
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = (T)d.small_w[t + q * EL];
#pragma unroll
    for (int q = 0; q < 4; ++q) xv[q] = x[(size_t)col[q] * K + c];
#pragma unroll
    for (int q = 0; q < 4; ++q) s += w[q] * xv[q];
  }
  for (; t < tend; t += EL) s += (T)d.small_w[t] * x[(size_t)d.small_col[t] * K + c];
  for (int o = K; o < 64; o <<= 1) s += __shfl_xor(s, o, 64);
  if (lane < K) u[(size_t)c * n_full + d.small_dof[sm]] = s;
}

// Shell corrections W[s][l][c] for all K columns: one workgroup per shell
// cell, CK columns per wave (K / CK waves, so a workgroup's waves always take
// the same branch and meet the same barriers). Lane l holds cell dof
// (lx, ly, lz).
//  - Cut cell: W = D_s u_s with the stored D_s = C_s - A_s. For K = 16 the
//    two waves each take 32 rows of W on the f64 MFMA (16x16x4): the A
//    operand is read from HBM (row l of the symmetric block is its column l),
//    the B operand is one u value per lane from LDS, so each byte of the
//    block and of u is read once. Otherwise each step reads one 512-byte row
//    segment against the wave's columns (a group of rows in registers, the
//    next group's loads in flight); the waves of one cell read the same rows
//    at the same time, so the block crosses HBM once.
//  - Empty cell: W = -A_s u_s with A_s = Kx My Mz + Mx Ky Mz +
//    Mx My (Kz + c Mz), by sum factorisation: one 1D contraction per axis
//    (x: mx u, kx u; y: my, ky; z: mz, kz + c mz), 4 P1^4 multiply-adds per
//    column instead of the dense block's 2 P1^6. The contraction along an
//    axis reads, per step, the P1^2 values at one coordinate of that axis;
//    each stage's input is stored with that axis slowest, so those values
//    are contiguous in LDS and the read is free of bank conflicts (the P1
//    lanes that share one value get it by broadcast).
// The column index of W is fastest, so the touch pass reads the K columns
// of one entry as one contiguous segment.
typedef double v4d __attribute__((ext_vector_type(4)));

// LDS layout of the shell kernel's per-wave values: rows of 2 CK columns,
// each row P1 slabs of P1^2 dofs in stage order. Slab and row padding put
// the 64 lanes of an MFMA operand read (16 columns x 4 slabs) on every bank
// pair exactly four times; a stage read stays 16 contiguous doubles.
template <int P1, int CK>
struct ShellLds {
  static constexpr int QQ = P1 * P1;
  static constexpr int SLAB = QQ + (P1 == 4 ? 4 : 0);
  static constexpr int ROW = P1 * SLAB + 1;
  static constexpr int WST = 2 * CK * ROW + 8;
  static __device__ __forceinline__ int at(int h, int r, int q) { return h * WST + r * ROW + (q / QQ) * SLAB + q % QQ; }
};

template <typename T, int P1, int K>
__global__ __launch_bounds__(64 * (K < 8 ? 1 : K / 8)) void k_mfk_shell(cc_sumfact_desc d,
                                                                      const T* __restrict__ blocks,
                                                                      const T* __restrict__ u,
                                                                      T* __restrict__ Wout) {
  constexpr int QQ = P1 * P1;
  constexpr int PP = QQ * P1;
  constexpr int CK = K < 8 ? K : 8;
  constexpr int WPC = K / CK;
  static_assert(PP <= 64 && WPC * CK == K, "one lane per cell dof, whole column groups");
  // Per wave: rows [0, CK) u of its columns, x slowest ((x*P1 + z)*P1 + y),
  // later the first stage value of each column; [CK, 2 CK) the second.
  using L = ShellLds<P1, CK>;
  __shared__ T su[WPC * L::WST];
  __shared__ T sf[6][QQ];  // mx kx my ky mz kz of the cell
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int c0 = wv * CK;
  const int s = blockIdx.x;
  const size_t n_full = (size_t)d.nb[0] * d.nb[1] * d.nb[2];
  const int ex = d.corr_cell[3 * s + 0];
  const int ey = d.corr_cell[3 * s + 1];
  const int ez = d.corr_cell[3 * s + 2];
  if (lane < PP) {
    const int lz = lane / QQ, ly = (lane / P1) % P1, lx = lane % P1;
    const int gx = wrap_or_invalid(ex + lx, d.nb[0], d.periodic[0]);
    const int gy = wrap_or_invalid(ey + ly, d.nb[1], d.periodic[1]);
    const int gz = wrap_or_invalid(ez + lz, d.nb[2], d.periodic[2]);
    const size_t g = ((size_t)gz * d.nb[1] + gy) * d.nb[0] + gx;
    const int q = (lx * P1 + lz) * P1 + ly;
#pragma unroll
    for (int c = 0; c < CK; ++c) su[L::at(wv, c, q)] = u[(size_t)(c0 + c) * n_full + g];
  }
  if (wv == 0) {
    for (int i = lane; i < QQ; i += 64) {
      sf[0][i] = (T)d.elem_m[0][(size_t)ex * QQ + i];
      sf[1][i] = (T)d.elem_k[0][(size_t)ex * QQ + i];
      sf[2][i] = (T)d.elem_m[1][(size_t)ey * QQ + i];
      sf[3][i] = (T)d.elem_k[1][(size_t)ey * QQ + i];
      sf[4][i] = (T)d.elem_m[2][(size_t)ez * QQ + i];
      sf[5][i] = (T)d.elem_k[2][(size_t)ez * QQ + i];
    }
  }
  __syncthreads();
  const bool act = lane < PP;
  const int l = act ? lane : 0;
  const int lz = l / QQ, ly = (l / P1) % P1, lx = l % P1;
  T w[CK];
#pragma unroll
  for (int c = 0; c < CK; ++c) w[c] = 0;
  const int blk = d.corr_block[s];
  if constexpr (K == 16 && P1 == 4 && std::is_same<T, double>::value) {
    if (blk >= 0) {
      // W rows [32 wv, 32 wv + 32) = two 16-row tiles. Step t covers block
      // rows j = 4t + kk (kk = lane / 16): A[i][kk] = D[16 m + i][j] is read
      // as D[j][16 m + i]; B[kk][c] = u_j of column c = lane % 16, which sits
      // in wave c / 8, row c % 8, at stage-order position 16 kk + t.
      const int i16 = lane & 15, kk = lane >> 4;
      const double* __restrict__ Db = blocks + (size_t)blk * PP * PP + (size_t)kk * PP + 32 * wv + i16;
      double a0[16], a1[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        a0[t] = Db[(size_t)(4 * t) * PP];
        a1[t] = Db[(size_t)(4 * t) * PP + 16];
      }
      v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
      const int bbase = L::at(i16 >> 3, i16 & 7, 16 * kk);
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const double bv = su[bbase + t];
        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[t], bv, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[t], bv, acc1, 0, 0, 0);
      }
      // f64 16x16 tile layout (unlike the f32 ones): register r of lane
      // (i16, kk) holds row kk + 4 r of column i16.
      const size_t row0 = (size_t)s * PP + 32 * wv + kk;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Wout[(row0 + 4 * r) * K + i16] = acc0[r];
        Wout[(row0 + 16 + 4 * r) * K + i16] = acc1[r];
      }
      return;
    }
  }
  if constexpr (K == 16 && P1 == 4 && std::is_same<T, float>::value) {
    if (blk >= 0) {
      // The same tiles on the f32 MFMA 16x16x4, whose A and B operands sit
      // as the f64 one's; its D tile differs: register r of lane (i16, kk)
      // holds row 4 kk + r of column i16.
      const int i16 = lane & 15, kk = lane >> 4;
      const float* __restrict__ Db = blocks + (size_t)blk * PP * PP + (size_t)kk * PP + 32 * wv + i16;
      float a0[16], a1[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        a0[t] = Db[(size_t)(4 * t) * PP];
        a1[t] = Db[(size_t)(4 * t) * PP + 16];
      }
      typedef float v4f __attribute__((ext_vector_type(4)));
      v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
      const int bbase = L::at(i16 >> 3, i16 & 7, 16 * kk);
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const float bv = su[bbase + t];
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[t], bv, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[t], bv, acc1, 0, 0, 0);
      }
      const size_t row0 = (size_t)s * PP + 32 * wv + 4 * kk;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Wout[(row0 + r) * K + i16] = acc0[r];
        Wout[(row0 + 16 + r) * K + i16] = acc1[r];
      }
      return;
    }
  }
  if (blk >= 0) {
    // Rows in groups of G, the next group's loads in flight meanwhile.
    constexpr int G = PP % 8 == 0 ? 8 : QQ;
    const T* __restrict__ Db = blocks + (size_t)blk * PP * PP + l;
    T cur[G], nxt[G];
#pragma unroll
    for (int r = 0; r < G; ++r) cur[r] = Db[(size_t)r * PP];
#pragma unroll 1
    for (int j0 = 0; j0 < PP; j0 += G) {
      if (j0 + G < PP) {
#pragma unroll
        for (int r = 0; r < G; ++r) nxt[r] = Db[(size_t)(j0 + G + r) * PP];
      }
#pragma unroll
      for (int r = 0; r < G; ++r) {
        const int j = j0 + r;  // (jz, jy, jx) = (j / QQ, j / P1 % P1, j % P1)
        const int q = ((j % P1) * P1 + j / QQ) * P1 + (j / P1) % P1;
#pragma unroll
        for (int c = 0; c < CK; ++c) w[c] += cur[r] * su[L::at(wv, c, q)];
      }
#pragma unroll
      for (int r = 0; r < G; ++r) cur[r] = nxt[r];
    }
  } else {
    const T cm = (T)d.mass_coef;
    // x: a = mx u, b = kx u, stored y slowest.
    T a[CK], b[CK], am[CK];
#pragma unroll
    for (int c = 0; c < CK; ++c) a[c] = b[c] = 0;
#pragma unroll
    for (int jx = 0; jx < P1; ++jx) {
      const T mx = sf[0][lx * P1 + jx], kx = sf[1][lx * P1 + jx];
      const int q = (jx * P1 + lz) * P1 + ly;
#pragma unroll
      for (int c = 0; c < CK; ++c) {
        const T uj = su[L::at(wv, c, q)];
        a[c] += mx * uj;
        b[c] += kx * uj;
      }
    }
    __syncthreads();
    if (act) {
      const int q = (ly * P1 + lz) * P1 + lx;
#pragma unroll
      for (int c = 0; c < CK; ++c) {
        su[L::at(wv, c, q)] = a[c];
        su[L::at(wv, CK + c, q)] = b[c];
      }
    }
    __syncthreads();
    // y: ky a + my b (summed, both meet mz) and my a; stored z slowest.
#pragma unroll
    for (int c = 0; c < CK; ++c) a[c] = b[c] = am[c] = 0;
#pragma unroll
    for (int jy = 0; jy < P1; ++jy) {
      const T my = sf[2][ly * P1 + jy], ky = sf[3][ly * P1 + jy];
      const int q = (jy * P1 + lz) * P1 + lx;
#pragma unroll
      for (int c = 0; c < CK; ++c) {
        const T va = su[L::at(wv, c, q)], vb = su[L::at(wv, CK + c, q)];
        am[c] += my * va;
        a[c] += ky * va;
        b[c] += my * vb;
      }
    }
    __syncthreads();
    if (act) {
      const int q = (lz * P1 + ly) * P1 + lx;
#pragma unroll
      for (int c = 0; c < CK; ++c) {
        su[L::at(wv, c, q)] = a[c] + b[c];
        su[L::at(wv, CK + c, q)] = am[c];
      }
    }
    __syncthreads();
    // z: -(mz (ky mx + my kx) u + (kz + c mz) my mx u).
#pragma unroll
    for (int jz = 0; jz < P1; ++jz) {
      const T mz = sf[4][lz * P1 + jz];
      const T kzc = fma_t(cm, mz, sf[5][lz * P1 + jz]);
      const int q = (jz * P1 + ly) * P1 + lx;
      // Explicit roundings, so every K gives each column the same result.
#pragma unroll
      for (int c = 0; c < CK; ++c) w[c] -= fma_t(mz, su[L::at(wv, c, q)], mul_t(kzc, su[L::at(wv, CK + c, q)]));
    }
  }
  if (act) {
#pragma unroll
    for (int c = 0; c < CK; ++c) Wout[((size_t)s * PP + l) * K + c0 + c] = w[c];
  }
}

// v[c][g] += sum of the shell outputs of dof g: lane (entry slot, column),
// 64 / K entry slots per touched dof, summed by a butterfly.
template <typename T, int K>
__global__ __launch_bounds__(256) void k_mfk_touch(cc_sumfact_desc d, const T* __restrict__ Win,
                                                   T* __restrict__ v, int PP) {
  constexpr int EL = 64 / K;
  const int lane = threadIdx.x & 63;
  const int c = lane % K, el = lane / K;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const size_t n_full = (size_t)d.nb[0] * d.nb[1] * d.nb[2];
  for (int t = wave; t < d.n_touch; t += nwaves) {
    T s = 0;
    const int iend = d.touch_ptr[t + 1];
    int i = d.touch_ptr[t] + el;
    // Four entries' indices, then their values, in flight at once; the sum
    // keeps the one-entry-at-a-time order.
    for (; i + 3 * EL < iend; i += 4 * EL) {
      const int a0 = d.touch_idx[i], a1 = d.touch_idx[i + EL];
      const int a2 = d.touch_idx[i + 2 * EL], a3 = d.touch_idx[i + 3 * EL];
      // Win[cell][local dof][K]: the K columns of one entry are one
      // contiguous 8 K-byte segment.
      const T w0 = Win[(size_t)a0 * K + c], w1 = Win[(size_t)a1 * K + c];
      const T w2 = Win[(size_t)a2 * K + c], w3 = Win[(size_t)a3 * K + c];
      s += w0;
      s += w1;
      s += w2;
      s += w3;
    }
    for (; i < iend; i += EL) s += Win[(size_t)d.touch_idx[i] * K + c];
    for (int o = K; o < 64; o <<= 1) s += __shfl_xor(s, o, 64);
    if (lane < K) v[(size_t)c * n_full + d.touch_dof[t]] += s;
  }
}

// extra[j][c] = sum_b E[b, j] v[c][small dof b] for the unknowns with E^T entries.
template <typename T, int K>
__global__ __launch_bounds__(256) void k_mfk_et_extra(cc_sumfact_desc d, const T* __restrict__ v,
                                                      T* __restrict__ extra) {
  constexpr int EL = 64 / K;
  const int lane = threadIdx.x & 63;
  const int c = lane % K, el = lane / K;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  const size_t n_full = (size_t)d.nb[0] * d.nb[1] * d.nb[2];
  for (int i = wave; i < d.n_et_rows; i += nwaves) {
    const int j = d.et_rows[i];
    T s = 0;
    const int tend = d.et_ptr[j + 1];
    int t = d.et_ptr[j] + el;
    // Four entries' index chains in flight at once, summed in entry order.
    for (; t + 3 * EL < tend; t += 4 * EL) {
      int b[4];
      T w[4], vv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q] = d.et_small[t + q * EL];
#pragma unroll
      for (int q = 0; q < 4; ++q) w[q] = (T)d.et_w[t + q * EL];
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q] = d.small_dof[b[q]];
#pragma unroll
      for (int q = 0; q < 4; ++q) vv[q] = v[(size_t)c * n_full + b[q]];
#pragma unroll
      for (int q = 0; q < 4; ++q) s += w[q] * vv[q];
    }
    for (; t < tend; t += EL) s += (T)d.et_w[t] * v[(size_t)c * n_full + d.small_dof[d.et_small[t]]];
    for (int o = K; o < 64; o <<= 1) s += __shfl_xor(s, o, 64);
    if (lane < K) extra[(size_t)j * K + c] = s;
  }
}

// y = E^T v for K columns; partial[block][c] = sum_j x[j][c] y[j][c].
template <typename T, int K>
__global__ __launch_bounds__(256) void k_mfk_scatter(cc_sumfact_desc d, const T* __restrict__ v,
                                                     const T* __restrict__ extra, T* __restrict__ y,
                                                     const T* __restrict__ x, double* __restrict__ partial) {
  // Through an LDS tile of 1024 / K unknowns x K columns: the reads of v are
  // runs of one column's grid, the writes of y whole K-column rows.
  constexpr int TJ = 1024 / K;
  __shared__ T tile[TJ][K + 1];
  __shared__ double red[4 * K];
  const size_t n_full = (size_t)d.nb[0] * d.nb[1] * d.nb[2];
  const size_t n_red = (size_t)d.n_red;
  const size_t ntiles = (n_red + TJ - 1) / TJ;
  const int c = threadIdx.x % K;  // the column of every element this thread writes
  double dot = 0.0;
  for (size_t tb = blockIdx.x; tb < ntiles; tb += gridDim.x) {
    const size_t j0 = tb * TJ;
    __syncthreads();  // the previous tile's rows are written
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = threadIdx.x + 256 * q;
      const int cc = k / TJ, ji = k % TJ;
      const size_t j = j0 + ji;
      if (j < n_red) tile[ji][cc] = v[(size_t)cc * n_full + d.red2full[j]];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ji = (threadIdx.x + 256 * q) / K;
      const size_t j = j0 + ji;
      if (j < n_red) {
        const size_t e = j * K + c;
        T sv = tile[ji][c];
        if (d.et_ptr[j + 1] > d.et_ptr[j]) sv += extra[e];
        y[e] = sv;
        if (partial) dot += (double)x[e] * (double)sv;
      }
    }
  }
  if (partial) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = K; o < 64; o <<= 1) dot += __shfl_xor(dot, o, 64);
    if (lane < K) red[wave * K + lane] = dot;
    __syncthreads();
    if (threadIdx.x < K)
      partial[(size_t)blockIdx.x * K + threadIdx.x] =
          (red[threadIdx.x] + red[K + threadIdx.x]) + (red[2 * K + threadIdx.x] + red[3 * K + threadIdx.x]);
  }
}

// Per-block, per-column partials of ((b - y).(b - y), b.b): [block][2][K].
template <int K>
__global__ __launch_bounds__(256) void k_mfk_residual(int n, const double* __restrict__ b, const double* __restrict__ y,
                                                      double* __restrict__ partial) {
  __shared__ double red[2][4 * K];
  const size_t total = (size_t)n * K;
  const int c = threadIdx.x % K;
  double rr = 0.0, bb = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const double r = b[e] - y[e];
    rr += r * r;
    bb += b[e] * b[e];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = K; o < 64; o <<= 1) {
    rr += __shfl_xor(rr, o, 64);
    bb += __shfl_xor(bb, o, 64);
  }
  if (lane < K) {
    red[0][wave * K + lane] = rr;
    red[1][wave * K + lane] = bb;
  }
  __syncthreads();
  if (threadIdx.x < 2 * K) {
    const int q = threadIdx.x / K, cc = threadIdx.x % K;
    partial[((size_t)blockIdx.x * 2 + q) * K + cc] =
        (red[q][cc] + red[q][K + cc]) + (red[q][2 * K + cc] + red[q][3 * K + cc]);
  }
}

// R = B - R (R holds A X on entry) with per-block, per-column partials of
// R.R: partial[block][c]. The outer residual of the mixed solve.
template <int K>
__global__ __launch_bounds__(256) void k_mfk_sub_colsq(int n, const double* __restrict__ B, double* __restrict__ R,
                                                       double* __restrict__ partial) {
  __shared__ double red[4 * K];
  const size_t total = (size_t)n * K;
  double rr = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const double r = B[e] - R[e];
    R[e] = r;
    rr += r * r;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = K; o < 64; o <<= 1) rr += __shfl_xor(rr, o, 64);
  if (lane < K) red[wave * K + lane] = rr;
  __syncthreads();
  if (threadIdx.x < K)
    partial[(size_t)blockIdx.x * K + threadIdx.x] =
        (red[threadIdx.x] + red[K + threadIdx.x]) + (red[2 * K + threadIdx.x] + red[3 * K + threadIdx.x]);
}

// Precision conversions of a [n][K] block.
__global__ __launch_bounds__(256) void k_d2f(size_t count, const double* __restrict__ a, float* __restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}
__global__ __launch_bounds__(256) void k_f2d(size_t count, const float* __restrict__ a, double* __restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    b[i] = (double)a[i];
}

template <typename T>
struct MfBuf {
  T *u, *v, *W, *extra;
};
using MfBuffers = MfBuf<double>;

// The float32 view of K-column buffers carved for float64 (each array at
// the same offset, half of it used).
static MfBuf<float> as_f32(const MfBuffers& b) {
  return MfBuf<float>{reinterpret_cast<float*>(b.u), reinterpret_cast<float*>(b.v), reinterpret_cast<float*>(b.W),
                      reinterpret_cast<float*>(b.extra)};
}

static size_t mf_bytes(const cc_sumfact_desc* d) {
  const size_t n_full = (size_t)d->nb[0] * d->nb[1] * d->nb[2];
  const int P1 = d->degree + 1;
  const size_t pp = (size_t)P1 * P1 * P1;
  return 2 * n_full * sizeof(double) + (size_t)d->n_corr * pp * sizeof(double) + (size_t)d->n_red * sizeof(double) +
         4 * 256;
}

static MfBuffers carve_mf(const cc_sumfact_desc* d, char* base) {
  const size_t n_full = (size_t)d->nb[0] * d->nb[1] * d->nb[2];
  const int P1 = d->degree + 1;
  const size_t pp = (size_t)P1 * P1 * P1;
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  MfBuffers b;
  b.u = reinterpret_cast<double*>(base);
  b.v = reinterpret_cast<double*>(base + al(n_full * sizeof(double)));
  b.W = reinterpret_cast<double*>(base + 2 * al(n_full * sizeof(double)));
  b.extra = reinterpret_cast<double*>(base + 2 * al(n_full * sizeof(double)) +
                                      al((size_t)d->n_corr * pp * sizeof(double)));
  return b;
}

static int grid1d(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > (size_t)kMaxParts) g = kMaxParts;
  return (int)(g < 1 ? 1 : g);
}

// Blocks for kernels without per-block partials whose items (one per 256 /
// per_block threads) are latency-bound chains: one item per wave or team,
// no cap below the grid limit.
static int grid_items(size_t items, int per_block) {
  size_t g = (items + per_block - 1) / per_block;
  if (g > 0x7fffffffull) g = 0x7fffffffull;
  return (int)(g < 1 ? 1 : g);
}

template <int P1>
static int mf_apply_t(const cc_sumfact_desc& d, const MfBuffers& b, const double* x, double* y, double* partial,
                      int* nparts, const ColState* st, hipStream_t s) {
  const size_t n_full = (size_t)d.nb[0] * d.nb[1] * d.nb[2];
  hipLaunchKernelGGL(k_mf_gather, dim3(grid1d(n_full)), dim3(256), 0, s, d, x, b.u);
  if (d.n_small > 0) hipLaunchKernelGGL(k_mf_gather_small, dim3((d.n_small + 3) / 4), dim3(256), 0, s, d, x, b.u);
  tensor_launch<double, P1>(d, b.u, b.v, 1, s);
  if (d.n_corr > 0) {
    hipLaunchKernelGGL((k_mfk_shell<double, P1, 1>), dim3(d.n_corr), dim3(64), 0, s, d, d.blocks, b.u, b.W);
    hipLaunchKernelGGL(k_mf_touch, dim3(grid_items(d.n_touch, 16)), dim3(256), 0, s, d, b.W, b.v);
  }
  if (d.n_et_rows > 0)
    hipLaunchKernelGGL(k_mf_et_extra, dim3(grid_items(d.n_et_rows, 16)), dim3(256), 0, s, d, b.v, b.extra, st);
  const int gs = grid1d(d.n_red);
  hipLaunchKernelGGL(k_mf_scatter, dim3(gs), dim3(256), 0, s, d, b.v, b.extra, y, x, partial, st);
  if (nparts) *nparts = gs;
  return (int)hipGetLastError();
}

static int mf_apply(const cc_sumfact_desc& d, const MfBuffers& b, const double* x, double* y, double* partial,
                    int* nparts, const ColState* st, hipStream_t s) {
  switch (d.degree) {
    case 2: return mf_apply_t<3>(d, b, x, y, partial, nparts, st, s);
    case 3: return mf_apply_t<4>(d, b, x, y, partial, nparts, st, s);
    default: return CC_E_ARG;
  }
}

// ||b - A x|| / ||b|| (||b|| = 0 -> 1) with the matrix-free operator; the
// per-block partials are summed on the host in block order.
static int mf_true_residual(const cc_sumfact_desc& d, const MfBuffers& mb, const double* bvec, const double* X,
                            double* scratch, double* rpart, double* rel, hipStream_t s) {
  CC_CHECK(mf_apply(d, mb, X, scratch, nullptr, nullptr, nullptr, s));
  const int g = grid1d(d.n_red);
  hipLaunchKernelGGL(k_mf_residual, dim3(g), dim3(256), 0, s, d.n_red, bvec, scratch, rpart);
  CC_CHECK(hipGetLastError());
  std::vector<double> hp(2 * (size_t)g);
  CC_CHECK(hipMemcpyAsync(hp.data(), rpart, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  double rr = 0.0, bb = 0.0;
  for (int i = 0; i < g; ++i) {
    rr += hp[2 * i];
    bb += hp[2 * i + 1];
  }
  double bn = std::sqrt(bb);
  if (bn == 0.0) bn = 1.0;
  *rel = std::sqrt(rr) / bn;
  return 0;
}

// K-column buffers: u, v [K][n_full]; W [n_corr][K][P]; extra [n_red][K].
static size_t mfk_bytes(const cc_sumfact_desc* d, int K) {
  const size_t n_full = (size_t)d->nb[0] * d->nb[1] * d->nb[2];
  const int P1 = d->degree + 1;
  const size_t pp = (size_t)P1 * P1 * P1;
  return (2 * n_full + (size_t)d->n_corr * pp + (size_t)d->n_red) * K * sizeof(double) + 4 * 256;
}

static MfBuffers carve_mfk(const cc_sumfact_desc* d, int K, char* base) {
  const size_t n_full = (size_t)d->nb[0] * d->nb[1] * d->nb[2];
  const int P1 = d->degree + 1;
  const size_t pp = (size_t)P1 * P1 * P1;
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  MfBuffers b;
  size_t off = 0;
  b.u = reinterpret_cast<double*>(base + off);
  off += al(n_full * K * sizeof(double));
  b.v = reinterpret_cast<double*>(base + off);
  off += al(n_full * K * sizeof(double));
  b.W = reinterpret_cast<double*>(base + off);
  off += al((size_t)d->n_corr * pp * K * sizeof(double));
  b.extra = reinterpret_cast<double*>(base + off);
  return b;
}

// One K-column apply in value type T; blocks: the cut-cell corrections in T
// (the descriptor's float64 blocks, or a float32 copy).
// A second stream and two events per host thread and device, made on first
// use: the K-column apply runs its shell corrections there, beside the
// tensor product on the caller's stream. The two read the same u and write
// different arrays; the touch kernel waits for both. The tensor product is
// bound by its own issue rate and the corrections by memory, so they share
// the CUs well. Default priority; never destroyed (one per thread).
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

static SideStream* side_stream() {
  constexpr int kDevices = 16;
  thread_local SideStream per_dev[kDevices];
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kDevices) return nullptr;
  SideStream& ss = per_dev[dev];
  if (!ss.s) {
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreateWithFlags(&e0, hipEventDisableTiming) != hipSuccess) return nullptr;
    if (hipEventCreateWithFlags(&e1, hipEventDisableTiming) != hipSuccess) return nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return nullptr;
    ss.fork = e0;
    ss.join = e1;
    ss.s = st;
  }
  return &ss;
}

template <typename T, int P1, int K>
static int mfk_apply_t(const cc_sumfact_desc& d, const T* blocks, const MfBuf<T>& b, const T* x, T* y,
                       double* partial, int* nparts, hipStream_t s) {
  const size_t n_full = (size_t)d.nb[0] * d.nb[1] * d.nb[2];
  constexpr int PP = P1 * P1 * P1;
  // One block per 1024-element tile: no partials, so no cap on the grid.
  hipLaunchKernelGGL((k_mfk_gather<T, K>), dim3(grid_items(n_full * K, 1024)), dim3(256), 0, s, d, x, b.u);
  if (d.n_small > 0)
    hipLaunchKernelGGL((k_mfk_gather_small<T, K>), dim3((d.n_small + 3) / 4), dim3(256), 0, s, d, x, b.u);
  // The shell corrections (u -> W) beside the tensor product (u -> v).
  SideStream* ss = d.n_corr > 0 && K > 1 ? side_stream() : nullptr;
  const dim3 shell_grid(d.n_corr), shell_block(64 * (K < 8 ? 1 : K / 8));
  // The tensor product is queued first, so that its blocks are dispatched
  // before the corrections' many blocks take the CUs.
  if (ss) CC_CHECK(hipEventRecord(ss->fork, s));
  tensor_launch<T, P1>(d, b.u, b.v, K, s);
  if (ss) {
    CC_CHECK(hipStreamWaitEvent(ss->s, ss->fork, 0));
    hipLaunchKernelGGL((k_mfk_shell<T, P1, K>), shell_grid, shell_block, 0, ss->s, d, blocks, b.u, b.W);
    CC_CHECK(hipEventRecord(ss->join, ss->s));
  }
  if (d.n_corr > 0) {
    if (ss)
      CC_CHECK(hipStreamWaitEvent(s, ss->join, 0));
    else
      hipLaunchKernelGGL((k_mfk_shell<T, P1, K>), shell_grid, shell_block, 0, s, d, blocks, b.u, b.W);
    hipLaunchKernelGGL((k_mfk_touch<T, K>), dim3(grid_items(d.n_touch, 4)), dim3(256), 0, s, d, b.W, b.v, PP);
  }
  if (d.n_et_rows > 0)
    hipLaunchKernelGGL((k_mfk_et_extra<T, K>), dim3(grid_items(d.n_et_rows, 4)), dim3(256), 0, s, d, b.v, b.extra);
  const int gs = grid1d((size_t)d.n_red * K);
  hipLaunchKernelGGL((k_mfk_scatter<T, K>), dim3(gs), dim3(256), 0, s, d, b.v, b.extra, y, x, partial);
  if (nparts) *nparts = gs;
  return (int)hipGetLastError();
}

template <typename T>
static int mfk_apply_any(const cc_sumfact_desc& d, const T* blocks, int K, const MfBuf<T>& b, const T* x, T* y,
                         double* partial, int* nparts, hipStream_t s) {
#define CC_MFK(P1V)                                                                    \
  switch (K) {                                                                        \
    case 1: return mfk_apply_t<T, P1V, 1>(d, blocks, b, x, y, partial, nparts, s);   \
    case 2: return mfk_apply_t<T, P1V, 2>(d, blocks, b, x, y, partial, nparts, s);   \
    case 4: return mfk_apply_t<T, P1V, 4>(d, blocks, b, x, y, partial, nparts, s);   \
    case 8: return mfk_apply_t<T, P1V, 8>(d, blocks, b, x, y, partial, nparts, s);   \
    case 16: return mfk_apply_t<T, P1V, 16>(d, blocks, b, x, y, partial, nparts, s); \
    default: return CC_E_ARG;                                                         \
  }
  if (d.degree == 2) CC_MFK(3)
  if (d.degree == 3) CC_MFK(4)
#undef CC_MFK
  return CC_E_ARG;
}

static int mfk_apply(const cc_sumfact_desc& d, int K, const MfBuffers& b, const double* x, double* y,
                     double* partial, int* nparts, hipStream_t s) {
  return mfk_apply_any<double>(d, d.blocks, K, b, x, y, partial, nparts, s);
}

// ||b_c - (A x)_c|| / ||b_c|| for the K columns (||b_c|| = 0 -> 1); partials
// summed on the host in block order.
static int mfk_true_residual(const cc_sumfact_desc& d, int K, const MfBuffers& mb, const double* B, const double* X,
                             double* scratch, double* rpart, double* rel, hipStream_t s) {
  CC_CHECK(mfk_apply(d, K, mb, X, scratch, nullptr, nullptr, s));
  const int g = grid1d((size_t)d.n_red * K);
  switch (K) {
    case 1: hipLaunchKernelGGL((k_mfk_residual<1>), dim3(g), dim3(256), 0, s, d.n_red, B, scratch, rpart); break;
    case 2: hipLaunchKernelGGL((k_mfk_residual<2>), dim3(g), dim3(256), 0, s, d.n_red, B, scratch, rpart); break;
    case 4: hipLaunchKernelGGL((k_mfk_residual<4>), dim3(g), dim3(256), 0, s, d.n_red, B, scratch, rpart); break;
    case 8: hipLaunchKernelGGL((k_mfk_residual<8>), dim3(g), dim3(256), 0, s, d.n_red, B, scratch, rpart); break;
    case 16: hipLaunchKernelGGL((k_mfk_residual<16>), dim3(g), dim3(256), 0, s, d.n_red, B, scratch, rpart); break;
    default: return CC_E_ARG;
  }
  CC_CHECK(hipGetLastError());
  std::vector<double> hp((size_t)g * 2 * K);
  CC_CHECK(hipMemcpyAsync(hp.data(), rpart, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  CC_CHECK(hipStreamSynchronize(s));
  for (int c = 0; c < K; ++c) {
    double rr = 0.0, bb = 0.0;
    for (int i = 0; i < g; ++i) {
      rr += hp[((size_t)i * 2 + 0) * K + c];
      bb += hp[((size_t)i * 2 + 1) * K + c];
    }
    double bn = std::sqrt(bb);
    if (bn == 0.0) bn = 1.0;
    rel[c] = std::sqrt(rr) / bn;
  }
  return 0;
}

static bool desc_ok(const cc_sumfact_desc* d) {
  if (!d || (d->degree != 2 && d->degree != 3) || d->n_red <= 0) return false;
  for (int k = 0; k < 3; ++k) {
    if (d->nb[k] <= 0 || d->ne[k] <= 0 || !d->band_m[k] || !d->band_k[k] || !d->elem_m[k] || !d->elem_k[k])
      return false;
    if (d->periodic[k] ? d->nb[k] != d->ne[k] : d->nb[k] != d->ne[k] + d->degree) return false;
  }
  if (!d->red2full || !d->full2red || !d->full2small || !d->small_ptr || !d->et_ptr) return false;
  if (d->n_et_rows < 0 || (d->n_et_rows > 0 && !d->et_rows)) return false;
  if (d->n_small > 0 && (!d->small_col || !d->small_w || !d->small_dof || !d->et_small || !d->et_w)) return false;
  if (d->n_corr > 0 && (!d->corr_cell || !d->corr_block || !d->blocks || !d->touch_dof || !d->touch_ptr ||
                        !d->touch_idx))
    return false;
  return true;
}

}  // namespace cc

using namespace cc;

extern "C" {

size_t cc_sumfact_workspace_bytes(const cc_sumfact_desc* d, int32_t solve) {
  if (!desc_ok(d)) return 0;
  size_t b = mf_bytes(d) + 512;
  if (solve) b += cg_operator_bytes((size_t)d->n_red) + (size_t)kMaxParts * 2 * sizeof(double) + 512;
  return b;
}

int cc_sumfact_apply(const cc_sumfact_desc* d, const double* x, double* y, void* workspace, size_t workspace_bytes,
                     cc_stream_t stream) {
  if (!desc_ok(d) || !x || !y || !workspace) return CC_E_ARG;
  if (workspace_bytes < cc_sumfact_workspace_bytes(d, 0)) return CC_E_WORKSPACE;
  MfBuffers b = carve_mf(d, (char*)workspace);
  return mf_apply(*d, b, x, y, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int cc_pcg_sumfact(const cc_sumfact_desc* d, const double* bvec, double* x, double tol, int32_t maxiter,
                   int32_t preconditioner, int32_t check_every, void* workspace, size_t workspace_bytes,
                   cc_pcg_result* result, double* history, const int32_t* snap_iters, int32_t n_snap,
                   double* snapshots, cc_stream_t stream) {
  if (!desc_ok(d) || !bvec || !x || !workspace || !result || maxiter < 0) return CC_E_ARG;
  if (n_snap < 0 || (n_snap > 0 && (!snap_iters || !snapshots))) return CC_E_ARG;
  if (workspace_bytes < cc_sumfact_workspace_bytes(d, 1)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* base = (char*)workspace;
  MfBuffers mb = carve_mf(d, base);
  const size_t off = (mf_bytes(d) + 511) & ~size_t(255);
  double* rpart = reinterpret_cast<double*>(base + off);
  char* cgws = base + off + (((size_t)kMaxParts * 2 * sizeof(double) + 511) & ~size_t(255));
  const size_t cgb = workspace_bytes - (size_t)(cgws - base);
  const cc_sumfact_desc desc = *d;
  auto apply = [&](const double* P, double* Q, double* partial, int* gs, const ColState* st) {
    return mf_apply(desc, mb, P, Q, partial, gs, st, s);
  };
  auto residual = [&](const double* X, double* scratch, double* rel) {
    CC_CHECK(mf_apply(desc, mb, X, scratch, nullptr, nullptr, nullptr, s));
    const int g = grid1d(desc.n_red);
    hipLaunchKernelGGL(k_mf_residual, dim3(g), dim3(256), 0, s, desc.n_red, bvec, scratch, rpart);
    CC_CHECK(hipGetLastError());
    std::vector<double> hp(2 * (size_t)g);
    CC_CHECK(hipMemcpyAsync(hp.data(), rpart, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    CC_CHECK(hipStreamSynchronize(s));
    double rr = 0.0, bb = 0.0;
    for (int i = 0; i < g; ++i) {
      rr += hp[2 * i];
      bb += hp[2 * i + 1];
    }
    double bn = std::sqrt(bb);
    if (bn == 0.0) bn = 1.0;
    *rel = std::sqrt(rr) / bn;
    return 0;
  };
  return cg_operator_f64(desc.n_red, apply, preconditioner ? desc.diag : nullptr, bvec, x, tol, maxiter,
                         preconditioner, check_every, cgws, cgb, result, history, s, residual, snap_iters, n_snap,
                         snapshots);
}

int cc_fcg_sumfact(const cc_sumfact_desc* d, const cc_cnn_desc* net, const double* bvec, double* x, double tol,
                   int32_t maxiter, int32_t check_every, void* workspace, size_t workspace_bytes,
                   cc_pcg_result* result, double* history, cc_stream_t stream) {
  if (!desc_ok(d) || !net || !bvec || !x || !workspace || !result || maxiter < 0 || net->n_red != d->n_red)
    return CC_E_ARG;
  if (workspace_bytes < cc_fcg_sumfact_workspace_bytes(d, net)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* base = (char*)workspace;
  MfBuffers mb = carve_mf(d, base);
  const size_t off = (mf_bytes(d) + 511) & ~size_t(255);
  double* rpart = reinterpret_cast<double*>(base + off);
  char* fws = base + off + (((size_t)kMaxParts * 2 * sizeof(double) + 511) & ~size_t(255));
  const size_t fb = workspace_bytes - (size_t)(fws - base);
  const cc_sumfact_desc desc = *d;
  auto apply = [&](const double* P, double* Q, double* partial, int* gs) {
    return mf_apply(desc, mb, P, Q, partial, gs, nullptr, s);
  };
  auto residual = [&](const double* X, double* scratch, double* rel) {
    CC_CHECK(mf_apply(desc, mb, X, scratch, nullptr, nullptr, nullptr, s));
    const int g = grid1d(desc.n_red);
    hipLaunchKernelGGL(k_mf_residual, dim3(g), dim3(256), 0, s, desc.n_red, bvec, scratch, rpart);
    CC_CHECK(hipGetLastError());
    std::vector<double> hp(2 * (size_t)g);
    CC_CHECK(hipMemcpyAsync(hp.data(), rpart, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    CC_CHECK(hipStreamSynchronize(s));
    double rr = 0.0, bb = 0.0;
    for (int i = 0; i < g; ++i) {
      rr += hp[2 * i];
      bb += hp[2 * i + 1];
    }
    double bn = std::sqrt(bb);
    if (bn == 0.0) bn = 1.0;
    *rel = std::sqrt(rr) / bn;
    return 0;
  };
  return fcg_operator(desc.n_red, apply, *net, desc.diag, bvec, x, tol, maxiter, check_every, fws, fb, result,
                      history, s, residual);
}

static bool k_ok(int k) { return k == 1 || k == 2 || k == 4 || k == 8 || k == 16; }

size_t cc_sumfact_k_workspace_bytes(const cc_sumfact_desc* d, int32_t k, int32_t solve) {
  if (!desc_ok(d) || !k_ok(k)) return 0;
  size_t b = mfk_bytes(d, k) + 512;
  if (solve) b += cg_operator_multi_bytes((size_t)d->n_red, k) + (size_t)kMaxParts * 2 * k * sizeof(double) + 1024;
  return b;
}

int cc_sumfact_apply_k(const cc_sumfact_desc* d, int32_t k, const double* X, double* Y, void* workspace,
                       size_t workspace_bytes, cc_stream_t stream) {
  if (!desc_ok(d) || !k_ok(k) || !X || !Y || !workspace) return CC_E_ARG;
  if (workspace_bytes < cc_sumfact_k_workspace_bytes(d, k, 0)) return CC_E_WORKSPACE;
  MfBuffers b = carve_mfk(d, k, (char*)workspace);
  return mfk_apply(*d, k, b, X, Y, nullptr, nullptr, (hipStream_t)stream);
}

int cc_pcg_sumfact_k(const cc_sumfact_desc* d, int32_t k, const double* B, double* X, double tol, int32_t maxiter,
                     int32_t preconditioner, int32_t check_every, void* workspace, size_t workspace_bytes,
                     cc_pcg_result* results, cc_stream_t stream) {
  if (!desc_ok(d) || !k_ok(k) || !B || !X || !workspace || !results || maxiter < 0) return CC_E_ARG;
  if (workspace_bytes < cc_sumfact_k_workspace_bytes(d, k, 1)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* base = (char*)workspace;
  MfBuffers mb = carve_mfk(d, k, base);
  const size_t off = (mfk_bytes(d, k) + 511) & ~size_t(255);
  double* rpart = reinterpret_cast<double*>(base + off);
  char* cgws = base + off + (((size_t)kMaxParts * 2 * k * sizeof(double) + 511) & ~size_t(255));
  const size_t cgb = workspace_bytes - (size_t)(cgws - base);
  const cc_sumfact_desc desc = *d;
  const int K = k;
  auto apply = [&](const double* P, double* Q, double* partial, int* gs) {
    return mfk_apply(desc, K, mb, P, Q, partial, gs, s);
  };
  auto residual = [&](const double* Xv, double* scratch, double* rel) {
    return mfk_true_residual(desc, K, mb, B, Xv, scratch, rpart, rel, s);
  };
  return cg_operator_multi(desc.n_red, k, apply, preconditioner ? desc.diag : nullptr, B, X, tol, maxiter,
                           preconditioner, check_every, cgws, cgb, results, s, residual);
}

size_t cc_sumfact_mixed_k_workspace_bytes(const cc_sumfact_desc* d, int32_t k) {
  if (!desc_ok(d) || !k_ok(k)) return 0;
  return mfk_bytes(d, k) + 1024 + cg_operator_mixed_multi_bytes((size_t)d->n_red, k) + 512;
}

int cc_sumfact_apply_k_f32(const cc_sumfact_desc* d, const float* blocks32, int32_t k, const float* X, float* Y,
                           void* workspace, size_t workspace_bytes, cc_stream_t stream) {
  if (!desc_ok(d) || !k_ok(k) || !X || !Y || !workspace || (d->n_corr > 0 && !blocks32)) return CC_E_ARG;
  if (workspace_bytes < cc_sumfact_k_workspace_bytes(d, k, 0)) return CC_E_WORKSPACE;
  const MfBuffers b = carve_mfk(d, k, (char*)workspace);
  return mfk_apply_any<float>(*d, blocks32, k, as_f32(b), X, Y, nullptr, nullptr, (hipStream_t)stream);
}

int cc_pcg_sumfact_mixed_k(const cc_sumfact_desc* d, const float* blocks32, int32_t k, const double* B, double* X,
                           double tol, int32_t maxiter, double inner_tol, int32_t max_outer, int32_t preconditioner,
                           int32_t check_every, void* workspace, size_t workspace_bytes, cc_pcg_result* results,
                           cc_stream_t stream) {
  if (!desc_ok(d) || !k_ok(k) || !B || !X || !workspace || !results || maxiter < 0 || max_outer < 1 ||
      !(inner_tol > 0.0 && inner_tol < 1.0) || (d->n_corr > 0 && !blocks32))
    return CC_E_ARG;
  if (workspace_bytes < cc_sumfact_mixed_k_workspace_bytes(d, k)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* base = (char*)workspace;
  const MfBuffers mb = carve_mfk(d, k, base);
  const MfBuf<float> mb32 = as_f32(mb);
  const size_t off = (mfk_bytes(d, k) + 511) & ~size_t(255);
  char* cgws = base + off;
  const size_t cgb = workspace_bytes - off;
  const cc_sumfact_desc desc = *d;
  const int K = k;
  auto residual = [&](const double* Xv, double* R64, double* partial, int* gs) {
    CC_CHECK(mfk_apply(desc, K, mb, Xv, R64, nullptr, nullptr, s));
    const int g = grid1d((size_t)desc.n_red * K);
    switch (K) {
      case 1: hipLaunchKernelGGL((k_mfk_sub_colsq<1>), dim3(g), dim3(256), 0, s, desc.n_red, B, R64, partial); break;
      case 2: hipLaunchKernelGGL((k_mfk_sub_colsq<2>), dim3(g), dim3(256), 0, s, desc.n_red, B, R64, partial); break;
      case 4: hipLaunchKernelGGL((k_mfk_sub_colsq<4>), dim3(g), dim3(256), 0, s, desc.n_red, B, R64, partial); break;
      case 8: hipLaunchKernelGGL((k_mfk_sub_colsq<8>), dim3(g), dim3(256), 0, s, desc.n_red, B, R64, partial); break;
      default: hipLaunchKernelGGL((k_mfk_sub_colsq<16>), dim3(g), dim3(256), 0, s, desc.n_red, B, R64, partial);
    }
    *gs = g;
    return (int)hipGetLastError();
  };
  auto apply32 = [&](const float* P, float* Q, double* partial, int* gs) {
    return mfk_apply_any<float>(desc, blocks32, K, mb32, P, Q, partial, gs, s);
  };
  return cg_operator_mixed_multi(desc.n_red, k, residual, apply32, preconditioner ? desc.diag : nullptr, X, tol,
                                 maxiter, inner_tol, max_outer, preconditioner, check_every, cgws, cgb, results, s);
}

size_t cc_fcg_sumfact_ml_workspace_bytes(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml) {
  if (!desc_ok(d) || !ml || !ml_desc_ok(*ml)) return 0;
  return mf_bytes(d) + 1024 + (size_t)kMaxParts * 2 * sizeof(double) + fcg_operator_ml_bytes((size_t)d->n_red, *ml) +
         512;
}

int cc_fcg_sumfact_ml(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml, const double* bvec, double* x, double tol,
                      int32_t maxiter, int32_t check_every, void* workspace, size_t workspace_bytes,
                      cc_pcg_result* result, double* history, cc_stream_t stream) {
  if (!desc_ok(d) || !ml || !ml_desc_ok(*ml) || !bvec || !x || !workspace || !result || maxiter < 0 ||
      ml->net[0].n_red != d->n_red)
    return CC_E_ARG;
  if (workspace_bytes < cc_fcg_sumfact_ml_workspace_bytes(d, ml)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* base = (char*)workspace;
  MfBuffers mb = carve_mf(d, base);
  const size_t off = (mf_bytes(d) + 511) & ~size_t(255);
  double* rpart = reinterpret_cast<double*>(base + off);
  char* fws = base + off + (((size_t)kMaxParts * 2 * sizeof(double) + 511) & ~size_t(255));
  const size_t fb = workspace_bytes - (size_t)(fws - base);
  const cc_sumfact_desc desc = *d;
  auto apply = [&](const double* P, double* Q, double* partial, int* gs) {
    return mf_apply(desc, mb, P, Q, partial, gs, nullptr, s);
  };
  auto residual = [&](const double* X, double* scratch, double* rel) {
    return mf_true_residual(desc, mb, bvec, X, scratch, rpart, rel, s);
  };
  return fcg_operator_ml(desc.n_red, apply, *ml, desc.diag, bvec, x, tol, maxiter, check_every, fws, fb, result,
                         history, s, residual);
}

size_t cc_fcg_sumfact_ml_k_workspace_bytes(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml, int32_t k) {
  if (!desc_ok(d) || !ml || !ml_desc_ok(*ml) || !k_ok(k)) return 0;
  return mfk_bytes(d, k) + 512 + (size_t)kMaxParts * 2 * k * sizeof(double) + 512 +
         fcg_operator_ml_k_bytes((size_t)d->n_red, *ml, k) + 512;
}

int cc_fcg_sumfact_ml_k(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml, int32_t k, const double* B, double* X,
                        double tol, int32_t maxiter, int32_t check_every, void* workspace, size_t workspace_bytes,
                        cc_pcg_result* results, cc_stream_t stream) {
  if (!desc_ok(d) || !ml || !ml_desc_ok(*ml) || !k_ok(k) || !B || !X || !workspace || !results || maxiter < 0 ||
      ml->net[0].n_red != d->n_red)
    return CC_E_ARG;
  if (workspace_bytes < cc_fcg_sumfact_ml_k_workspace_bytes(d, ml, k)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* base = (char*)workspace;
  MfBuffers mb = carve_mfk(d, k, base);
  const size_t off = (mfk_bytes(d, k) + 511) & ~size_t(255);
  double* rpart = reinterpret_cast<double*>(base + off);
  char* fws = base + off + (((size_t)kMaxParts * 2 * k * sizeof(double) + 511) & ~size_t(255));
  const size_t fb = workspace_bytes - (size_t)(fws - base);
  const cc_sumfact_desc desc = *d;
  const int K = k;
  auto apply = [&](const double* P, double* Q, double* partial, int* gs) {
    return mfk_apply(desc, K, mb, P, Q, partial, gs, s);
  };
  auto residual = [&](const double* Xv, double* scratch, double* rel) {
    return mfk_true_residual(desc, K, mb, B, Xv, scratch, rpart, rel, s);
  };
  return fcg_operator_ml_k(desc.n_red, k, apply, *ml, desc.diag, B, X, tol, maxiter, check_every, fws, fb, results,
                           s, residual);
}

// Workspace of cc_fcg_sumfact_ml_mixed_k: the K-column operator buffers, the
// float32 direction and product blocks, the flexible CG's, the refinement's.
static size_t ml_mixed_parts(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml, int k, size_t* off32, size_t* offf,
                             size_t* offr) {
  auto al = [](size_t x) { return (x + 511) & ~size_t(255); };
  const size_t nk = (size_t)d->n_red * k;
  size_t b = al(mfk_bytes(d, k));
  *off32 = b;
  b += al(2 * nk * sizeof(float));
  *offf = b;
  b += al(fcg_operator_ml_k_bytes((size_t)d->n_red, *ml, k));
  *offr = b;
  b += al(refine_operator_multi_bytes((size_t)d->n_red, k));
  return b;
}

size_t cc_fcg_sumfact_ml_mixed_k_workspace_bytes(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml, int32_t k) {
  if (!desc_ok(d) || !ml || !ml_desc_ok(*ml) || !k_ok(k)) return 0;
  size_t a, b, c;
  return ml_mixed_parts(d, ml, k, &a, &b, &c);
}

int cc_fcg_sumfact_ml_mixed_k(const cc_sumfact_desc* d, const float* blocks32, const cc_mlcnn_desc* ml, int32_t k,
                              const double* B, double* X, double tol, int32_t maxiter, double inner_tol,
                              int32_t max_outer, int32_t check_every, void* workspace, size_t workspace_bytes,
                              cc_pcg_result* results, cc_stream_t stream) {
  if (!desc_ok(d) || !ml || !ml_desc_ok(*ml) || !k_ok(k) || !B || !X || !workspace || !results || maxiter < 0 ||
      ml->net[0].n_red != d->n_red || max_outer < 1 || !(inner_tol > 0.0 && inner_tol < 1.0) ||
      (d->n_corr > 0 && !blocks32))
    return CC_E_ARG;
  if (workspace_bytes < cc_fcg_sumfact_ml_mixed_k_workspace_bytes(d, ml, k)) return CC_E_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  char* base = (char*)workspace;
  size_t off32, offf, offr;
  const size_t total = ml_mixed_parts(d, ml, k, &off32, &offf, &offr);
  const MfBuffers mb = carve_mfk(d, k, base);
  const MfBuf<float> mb32 = as_f32(mb);
  const size_t nk = (size_t)d->n_red * k;
  float* P32 = reinterpret_cast<float*>(base + off32);
  float* Q32 = P32 + nk;
  char* fws = base + offf;
  const size_t fb = offr - offf;
  char* rws = base + offr;
  const size_t rb = total - offr;
  const cc_sumfact_desc desc = *d;
  const int K = k;
  const int gconv = grid1d(nk);
  // The flexible CG's operator: its float64 direction rounded to float32,
  // the float32 apply (with per-block partials of P32.Q32), the product
  // widened back.
  auto apply32 = [&](const double* P, double* Q, double* partial, int* gs) {
    hipLaunchKernelGGL(k_d2f, dim3(gconv), dim3(256), 0, s, nk, P, P32);
    CC_CHECK(mfk_apply_any<float>(desc, blocks32, K, mb32, P32, Q32, partial, gs, s));
    hipLaunchKernelGGL(k_f2d, dim3(gconv), dim3(256), 0, s, nk, Q32, Q);
    return (int)hipGetLastError();
  };
  // The inner solve reports recursive residuals only; the refinement takes
  // the true ones.
  auto no_residual = [&](const double*, double*, double* rel) {
    for (int c = 0; c < K; ++c) rel[c] = 0.0;
    return 0;
  };
  auto inner = [&](const double* Bs, double* Z, int imax, int* its) {
    std::vector<cc_pcg_result> r(K);
    CC_CHECK(fcg_operator_ml_k(desc.n_red, K, apply32, *ml, desc.diag, Bs, Z, inner_tol, imax, check_every, fws, fb,
                               r.data(), s, no_residual));
    for (int c = 0; c < K; ++c) its[c] = r[c].iterations;
    return 0;
  };
  auto residual = [&](const double* Xv, double* R64, double* partial, int* gs) {
    CC_CHECK(mfk_apply(desc, K, mb, Xv, R64, nullptr, nullptr, s));
    const int g = grid1d(nk);
    switch (K) {
      case 1: hipLaunchKernelGGL((k_mfk_sub_colsq<1>), dim3(g), dim3(256), 0, s, desc.n_red, B, R64, partial); break;
      case 2: hipLaunchKernelGGL((k_mfk_sub_colsq<2>), dim3(g), dim3(256), 0, s, desc.n_red, B, R64, partial); break;
      case 4: hipLaunchKernelGGL((k_mfk_sub_colsq<4>), dim3(g), dim3(256), 0, s, desc.n_red, B, R64, partial); break;
      case 8: hipLaunchKernelGGL((k_mfk_sub_colsq<8>), dim3(g), dim3(256), 0, s, desc.n_red, B, R64, partial); break;
      default: hipLaunchKernelGGL((k_mfk_sub_colsq<16>), dim3(g), dim3(256), 0, s, desc.n_red, B, R64, partial);
    }
    *gs = g;
    return (int)hipGetLastError();
  };
  return refine_operator_multi(desc.n_red, k, residual, inner, X, tol, maxiter, max_outer, rws, rb, results, s);
}

size_t cc_fcg_sumfact_workspace_bytes(const cc_sumfact_desc* d, const cc_cnn_desc* net) {
  if (!desc_ok(d) || !net) return 0;
  return mf_bytes(d) + 1024 + (size_t)kMaxParts * 2 * sizeof(double) +
         fcg_operator_bytes((size_t)d->n_red, (size_t)net->nx * net->ny * net->nz) + 512;
}

}  // extern "C"
