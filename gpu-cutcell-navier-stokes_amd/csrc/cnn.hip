// cnn.hip - masked 3x3x3 convolution layers for the learned preconditioner.
//
// Activations live on the background dof grid (nx, ny, nz; x fastest) as
// channel-last bfloat16 tensors [voxel][16]. A layer is
//     out[v][co] = mask[v] * act( sum_t sum_ci W[t][ci][co] in[v + off_t][ci] )
// with t over the 27 taps (dz, dy, dx in {-1, 0, 1}), zero padding on clamped
// axes and wrap-around on periodic ones. As a GEMM this is M = voxels,
// N = 16 output channels, K = 27 taps x 16 input channels: every tap is one
// 16x16x16 MFMA (bf16 inputs, fp32 accumulate) on a wave's 16 consecutive
// x-voxels, so a wave issues 27 MFMAs per 16 voxels.
//
// MFMA operand layout (v_mfma_f32_16x16x16_bf16, wave64), lane l:
//   A (16 x 16, M x K): a[i] = A[l % 16][4 (l / 16) + i]
//   B (16 x 16, K x N): b[i] = B[4 (l / 16) + i][l % 16]
//   D (16 x 16, M x N): d[i] = D[4 (l / 16) + i][l % 16]
// (checked on the device by cc_mfma_probe against a host product).
//
// Backward: the data gradient is the same layer with the taps reversed and
// W transposed (prepared by cc_cnn_flip_weights), the weight gradient
//     dW[t][ci][co] = sum_v in[v + off_t][ci] g[v][co]
// is a GEMM over voxels (M = ci, N = co, K = voxels) with per-block fp32
// partials reduced in a fixed order by a second kernel.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Round-to-nearest-even float -> bfloat16 (NaN kept quiet).
__device__ __host__ inline uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __host__ inline float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__device__ inline int nb_index(int i, int n, int periodic) {
  if (periodic) {
    i = i % n;
    return i < 0 ? i + n : i;
  }
  return (i >= 0 && i < n) ? i : -1;
}

struct GridArgs {
  int nx, ny, nz;
  int px, py, pz;  // periodic flags
};

// Forward (and data-gradient) layer; ACT: 0 none, 1 relu. gate (optional,
// bf16 [voxel][16]): multiply the result by (gate > 0), the ReLU derivative
// in backward mode.
// A block of 8 waves owns a 16 x 4 x 8 (x, y, z) tile of output voxels; wave
// w the 4 rows (y0 .. y0 + 3) of 16 x-consecutive voxels in plane z0 + w,
// one MFMA accumulator per row. The block stages the tile's 18 x 6 x 10
// input halo (all 16 channels, zero or wrapped outside the grid) and the
// layer's weights in LDS, both split by channel group: halo[kg][voxel][4]
// and wl[tap][kg][co][4], so that the 16 lanes of one k group read 16
// consecutive 8-byte operands (bank-conflict free; a [voxel][16] layout
// puts those lanes 32 bytes apart, a 4-way conflict). For each (dz, dx) a
// wave reads the A operands of the 6 halo rows its 4 output rows touch once,
// and each of the 3 dy taps' B operands once, for 12 MFMAs.
constexpr int kCx = 16, kCy = 4, kCz = 8;
constexpr int kHx = kCx + 2, kHy = kCy + 2, kHz = kCz + 2;
constexpr int kConvThreads = 64 * kCz;  // one wave per z-plane of the tile

// A block owns one (x, y) column of tiles and runs a run of work items
// (grid b of the batch, z-tile) in turn: items blockIdx.z * ipb ..
// min(nb * tiles_z, (blockIdx.z + 1) * ipb) - 1, item = b * tiles_z + z-tile.
// Grid b of in / gate / out / out_f32 starts b * bstride elements in (the
// mask is shared). The weights are staged once per block; the next item's
// halo is loaded from memory into registers while the current item's MFMAs
// run.
constexpr int kHaloChunks = kHz * kHy * kHx * 2;  // 16-byte chunks of a halo
constexpr int kHaloPerThread = (kHaloChunks + kConvThreads - 1) / kConvThreads;
// Elements of one channel-group plane of the halo: 4 channels of every halo
// voxel, plus 16 bytes so that the planes a 16-byte chunk is split over
// start in different banks.
constexpr int kHaloPlane = kHz * kHy * kHx * 4 + 8;

__device__ inline void halo_fetch(const GridArgs& g, const uint16_t* __restrict__ inb,
                                  const float* __restrict__ inb_f32, int x0, int y0, int z0,
                                  uint4 (&reg)[kHaloPerThread]) {
#pragma unroll
  for (int k = 0; k < kHaloPerThread; ++k) {
    const int i = threadIdx.x + k * kConvThreads;
    uint4 v = {0u, 0u, 0u, 0u};
    if (i < kHaloChunks) {
      const int half = i & 1, hv = i >> 1;
      const int hx = hv % kHx, hy = (hv / kHx) % kHy, hz = hv / (kHx * kHy);
      const int gx = nb_index(x0 - 1 + hx, g.nx, g.px);
      const int gy = nb_index(y0 - 1 + hy, g.ny, g.py);
      const int gz = nb_index(z0 - 1 + hz, g.nz, g.pz);
      if (gx >= 0 && gy >= 0 && gz >= 0) {
        const size_t gv = ((size_t)gz * g.ny + gy) * g.nx + gx;
        if (inb_f32) {
          if (half == 0) v.x = f2bf(inb_f32[gv]);
        } else {
          v = *reinterpret_cast<const uint4*>(inb + gv * 16 + 8 * half);
        }
      }
    }
    reg[k] = v;
  }
}

// Chunk i (channels 8 h .. 8 h + 7 of halo voxel hv, h = i & 1) goes to the
// channel-group planes 2 h and 2 h + 1.
__device__ inline void halo_store(uint16_t* halo, const uint4 (&reg)[kHaloPerThread]) {
#pragma unroll
  for (int k = 0; k < kHaloPerThread; ++k) {
    const int i = threadIdx.x + k * kConvThreads;
    if (i < kHaloChunks) {
      const int hv = i >> 1, h = i & 1;
      *reinterpret_cast<uint2*>(&halo[(2 * h) * kHaloPlane + hv * 4]) = make_uint2(reg[k].x, reg[k].y);
      *reinterpret_cast<uint2*>(&halo[(2 * h + 1) * kHaloPlane + hv * 4]) = make_uint2(reg[k].z, reg[k].w);
    }
  }
}

// in_f32 (optional, replaces in): a single-channel float32 input grid (the
// other 15 input channels are zero), batch stride bstride / 16. out_ch0:
// out_f32 receives channel 0 only, as a single-channel grid (batch stride
// bstride / 16).
template <int ACT>
__global__ __launch_bounds__(kConvThreads) void k_conv3(GridArgs g, const uint16_t* __restrict__ in,
                                                        const float* __restrict__ in_f32,
                                                        const uint16_t* __restrict__ W,
                                                        const uint8_t* __restrict__ mask,
                                                        const uint16_t* __restrict__ gate, uint16_t* __restrict__ out,
                                                        float* __restrict__ out_f32, int out_ch0, int tiles_z,
                                                        size_t bstride, int nb, int ipb) {
  __shared__ __align__(16) uint16_t halo[4 * kHaloPlane];
  __shared__ __align__(16) uint16_t wl[27 * 16 * 16];
  // Row stride 20 floats: the transposing writes (4-byte) and reads (16-byte)
  // of a wave are both bank-conflict free.
  __shared__ __align__(16) float tile[kCz][16][20];
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * kCx, y0 = blockIdx.y * kCy;
  const int it0 = blockIdx.z * ipb;
  const int it1 = min(nb * tiles_z, it0 + ipb);
  // Weights W[t][ci][co] -> wl[t][ci / 4][co][ci % 4]: 16-byte loads of 8
  // output channels, scattered into LDS.
  for (int i = tid; i < 27 * 16 * 2; i += blockDim.x) {
    const int half = i & 1, tc = i >> 1;  // tc = t * 16 + ci
    const int t = tc / 16, ci = tc % 16;
    const uint4 v = *reinterpret_cast<const uint4*>(W + (size_t)tc * 16 + 8 * half);
    const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int co = 8 * half + k;
      wl[((t * 4 + ci / 4) * 16 + co) * 4 + ci % 4] = (uint16_t)(w4[k >> 1] >> (16 * (k & 1)));
    }
  }
  const int lane = tid & 63;
  const int wv = tid >> 6;    // z-plane of the tile
  const int row = lane % 16;  // voxel within a row of 16 (A row / D col index)
  const int kg = lane / 16;   // k group: channels 4 kg .. 4 kg + 3
  uint4 reg[kHaloPerThread];
  const size_t bstride1 = bstride / 16;  // batch stride of single-channel grids
  if (it0 < it1) {
    const int b = it0 / tiles_z;
    halo_fetch(g, in_f32 ? nullptr : in + (size_t)b * bstride, in_f32 ? in_f32 + (size_t)b * bstride1 : nullptr, x0,
               y0, (it0 - b * tiles_z) * kCz, reg);
  }
  for (int it = it0; it < it1; ++it) {
    const int b = it / tiles_z;
    const int z0 = (it - b * tiles_z) * kCz;
    const size_t off = (size_t)b * bstride;
    // Every wave is done with the previous item's halo (and the weights are
    // staged) before the halo is overwritten; then it is complete.
    __syncthreads();
    halo_store(halo, reg);
    __syncthreads();
    // The next item's halo is in flight while this item's MFMAs run.
    if (it + 1 < it1) {
      const int bn = (it + 1) / tiles_z;
      halo_fetch(g, in_f32 ? nullptr : in + (size_t)bn * bstride, in_f32 ? in_f32 + (size_t)bn * bstride1 : nullptr,
                 x0, y0, (it + 1 - bn * tiles_z) * kCz, reg);
    }
    v4f acc[kCy];
#pragma unroll
    for (int r = 0; r < kCy; ++r) acc[r] = v4f{0.f, 0.f, 0.f, 0.f};
    // dz is not unrolled: one plane's operands in flight at a time keeps the
    // kernel within the registers of 4 waves per SIMD.
#pragma unroll 1
    for (int dz = 0; dz < 3; ++dz) {
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        v4s a[kHy];
#pragma unroll
        for (int hy = 0; hy < kHy; ++hy) {
          const int hv = ((wv + dz) * kHy + hy) * kHx + (row + dx);
          a[hy] = *reinterpret_cast<const v4s*>(&halo[kg * kHaloPlane + hv * 4]);
        }
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
          const int t = dz * 9 + dy * 3 + dx;
          const v4s bw = *reinterpret_cast<const v4s*>(&wl[((t * 4 + kg) * 16 + row) * 4]);
#pragma unroll
          for (int r = 0; r < kCy; ++r) acc[r] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[r + dy], bw, acc[r], 0, 0, 0);
        }
      }
    }
    // acc[r][i] = D[voxel 4 kg + i][co = row] of row r; transpose each row
    // through this wave's own slice of tile (LDS accesses of one wave
    // complete in order, and no other wave touches the slice) so that each
    // lane stores 4 consecutive channels of one voxel.
    const int vx = x0 + row, vz = z0 + wv;
#pragma unroll
    for (int r = 0; r < kCy; ++r) {
#pragma unroll
      for (int i = 0; i < 4; ++i) tile[wv][4 * kg + i][row] = acc[r][i];
      const int vy = y0 + r;
      if (vx < g.nx && vy < g.ny && vz < g.nz) {
        const size_t v = ((size_t)vz * g.ny + vy) * g.nx + vx;
        const bool live = mask[v] != 0;
        uint16_t h[4];
        float f[4];
        const float4 tv = *reinterpret_cast<const float4*>(&tile[wv][row][4 * kg]);
        const float tf[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float val = tf[i];
          if (ACT == 1) val = val > 0.f ? val : 0.f;
          if (gate) val = bf2f(gate[off + v * 16 + 4 * kg + i]) > 0.f ? val : 0.f;
          if (!live) val = 0.f;
          f[i] = val;
          h[i] = f2bf(val);
        }
        if (out) {
          uint2 raw;
          raw.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
          raw.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
          *reinterpret_cast<uint2*>(out + off + v * 16 + 4 * kg) = raw;
        }
        if (out_f32) {
          if (out_ch0) {
            if (kg == 0) out_f32[(size_t)b * bstride1 + v] = f[0];
          } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) out_f32[off + v * 16 + 4 * kg + i] = f[i];
          }
        }
      }
    }
  }
}

// The network's first layer (one input channel, u as single-channel float32
// grids): out[v][co] = mask[v] sum_t W[t][0][co] u[v + off_t]. With a single
// input channel the 27 taps are the GEMM's K dimension: two MFMAs (taps
// 0..15 and 16..31, those past 26 zero) per row of 16 x-voxels, in place of
// k_conv3's 27. The tile, the halo staging (channel 0 of the bf16 halo
// planes) and the output transpose are k_conv3's.
__global__ __launch_bounds__(kConvThreads) void k_conv3_first(GridArgs g, const float* __restrict__ u,
                                                              const uint16_t* __restrict__ W,
                                                              const uint8_t* __restrict__ mask,
                                                              uint16_t* __restrict__ out, int tiles_z, size_t bstride,
                                                              int nb, int ipb) {
  __shared__ __align__(16) uint16_t halo[4 * kHaloPlane];
  __shared__ __align__(16) float tile[kCz][16][20];
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * kCx, y0 = blockIdx.y * kCy;
  const int it0 = blockIdx.z * ipb;
  const int it1 = min(nb * tiles_z, it0 + ipb);
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const int row = lane % 16;
  const int kg = lane / 16;
  // B operands: W[t][0][co = row] for t = 16 j + 4 kg + i (zero past 26).
  v4s bw[2];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int t = 16 * j + 4 * kg + i;
      bw[j][i] = t < 27 ? (short)W[(size_t)t * 256 + row] : (short)0;
    }
  uint4 reg[kHaloPerThread];
  const size_t bstride1 = bstride / 16;
  if (it0 < it1) {
    const int b = it0 / tiles_z;
    halo_fetch(g, nullptr, u + (size_t)b * bstride1, x0, y0, (it0 - b * tiles_z) * kCz, reg);
  }
  for (int it = it0; it < it1; ++it) {
    const int b = it / tiles_z;
    const int z0 = (it - b * tiles_z) * kCz;
    const size_t off = (size_t)b * bstride;
    __syncthreads();
    halo_store(halo, reg);
    __syncthreads();
    if (it + 1 < it1) {
      const int bn = (it + 1) / tiles_z;
      halo_fetch(g, nullptr, u + (size_t)bn * bstride1, x0, y0, (it + 1 - bn * tiles_z) * kCz, reg);
    }
    const int vx = x0 + row, vz = z0 + wv;
#pragma unroll
    for (int r = 0; r < kCy; ++r) {
      v4f acc = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        v4s a;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int t = 16 * j + 4 * kg + i;
          const int tt = t < 27 ? t : 26;
          const int hv = ((wv + tt / 9) * kHy + (r + (tt / 3) % 3)) * kHx + (row + tt % 3);
          const short v = (short)halo[hv * 4];  // channel 0 lives in plane 0, slot 0
          a[i] = t < 27 ? v : (short)0;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw[j], acc, 0, 0, 0);
      }
      // acc[i] = D[voxel 4 kg + i][co = row]: transpose through this wave's
      // slice so that each lane holds 4 channels of voxel `row`.
#pragma unroll
      for (int i = 0; i < 4; ++i) tile[wv][4 * kg + i][row] = acc[i];
      const int vy = y0 + r;
      if (vx < g.nx && vy < g.ny && vz < g.nz) {
        const size_t v = ((size_t)vz * g.ny + vy) * g.nx + vx;
        const bool live = mask[v] != 0;
        const float4 tv = *reinterpret_cast<const float4*>(&tile[wv][row][4 * kg]);
        const float tf[4] = {tv.x, tv.y, tv.z, tv.w};
        uint16_t h[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) h[i] = f2bf(live ? tf[i] : 0.f);
        uint2 raw;
        raw.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
        raw.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
        *reinterpret_cast<uint2*>(out + off + v * 16 + 4 * kg) = raw;
      }
    }
  }
}

static dim3 conv_grid(int nx, int ny, int nz, int nb = 1, int ipb = 1) {
  const int items = ((nz + kCz - 1) / kCz) * nb;
  return dim3((unsigned)((nx + kCx - 1) / kCx), (unsigned)((ny + kCy - 1) / kCy), (unsigned)((items + ipb - 1) / ipb));
}
static int conv_tiles_z(int nz) { return (nz + kCz - 1) / kCz; }

// Work items (grid of the batch, z-tile) per block: about kConvBlocks blocks
// in all (two fit a CU by LDS), each running its items in turn with the
// next item's halo loaded during the current one's MFMAs. g_conv_ipb > 0
// (cc_cnn_set_config) fixes it instead.
constexpr int kConvBlocks = 1024;
static int g_conv_ipb = 0;
static int conv_ipb(int nx, int ny, int nz, int nb) {
  const long long xy = (long long)((nx + kCx - 1) / kCx) * ((ny + kCy - 1) / kCy);
  const long long items = (long long)((nz + kCz - 1) / kCz) * nb;
  long long ipb = g_conv_ipb > 0 ? g_conv_ipb : (xy * items + kConvBlocks - 1) / kConvBlocks;
  if (ipb < 1) ipb = 1;
  if (ipb > items) ipb = items;
  return (int)ipb;
}

// Weight gradient partials: part[block][t][ci][co] = sum over the block's
// voxel rows of in[v + off_t][ci] * gout[v][co].
__global__ __launch_bounds__(256) void k_conv3_wgrad(GridArgs g, const uint16_t* __restrict__ in,
                                                     const uint16_t* __restrict__ gout, float* __restrict__ part,
                                                     int rows_per_block) {
  __shared__ float red[27 * 256];  // 27 x 16 x 16 fp32 (27 KB), waves added in order
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int m = lane % 16;  // ci (A row) / co (B col)
  const int kg = lane / 16;
  const int xt = (g.nx + 15) / 16;
  const long long n_rows = (long long)xt * g.ny * g.nz;
  v4f acc[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  const long long r0 = (long long)blockIdx.x * rows_per_block;
  for (long long r = r0 + wv; r < r0 + rows_per_block && r < n_rows; r += 4) {
    const int xb = (int)(r % xt) * 16;
    const int y = (int)((r / xt) % g.ny);
    const int z = (int)(r / ((long long)xt * g.ny));
    // B operand: gout[voxel xb + 4 kg + i][co = m]. Loads are unconditional
    // (a clamped in-grid address, zeroed by a select) so they all overlap.
    v4s bb;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int x = xb + 4 * kg + i;
      const bool ok = x < g.nx;
      const short v = (short)gout[(((size_t)z * g.ny + y) * g.nx + (ok ? x : 0)) * 16 + m];
      bb[i] = ok ? v : (short)0;
    }
    int zi[3], yi[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      zi[k] = nb_index(z + k - 1, g.nz, g.pz);
      yi[k] = nb_index(y + k - 1, g.ny, g.py);
    }
    int xi[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int x = xb + 4 * kg + i;
#pragma unroll
      for (int k = 0; k < 3; ++k) xi[i][k] = (x < g.nx) ? nb_index(x + k - 1, g.nx, g.px) : -1;
    }
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      const int zz = zi[t / 9], yy = yi[(t / 3) % 3];
      v4s aa;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int xx = xi[i][t % 3];
        const bool ok = zz >= 0 && yy >= 0 && xx >= 0;
        const short v = (short)in[((ok ? ((size_t)zz * g.ny + yy) * g.nx + xx : 0)) * 16 + m];
        aa[i] = ok ? v : (short)0;
      }
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aa, bb, acc[t], 0, 0, 0);
    }
  }
  // acc[t][i] = dW[t][ci = 4 kg + i][co = m]; add the 4 waves in a fixed order.
  for (int w = 0; w < 4; ++w) {
    if (wv == w) {
#pragma unroll
      for (int t = 0; t < 27; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int e = t * 256 + (4 * kg + i) * 16 + m;
          red[e] = (w == 0 ? 0.f : red[e]) + acc[t][i];
        }
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < 27 * 256; e += blockDim.x) part[(size_t)blockIdx.x * 27 * 256 + e] = red[e];
}

// out[e] (+)= scale * sum_b part[b][e]: a block covers 32 outputs with 8
// lanes each; lane q sums parts q, q + 8, ... (two running sums), and the 8
// lane sums are combined in a fixed order.
__global__ __launch_bounds__(256) void k_reduce_parts(const float* __restrict__ part, int nparts, int len,
                                                      float* __restrict__ out, float scale, int accumulate) {
  __shared__ double acc[8][32];
  const int o = threadIdx.x & 31;
  const int q = threadIdx.x >> 5;
  const int e = blockIdx.x * 32 + o;
  double s0 = 0.0, s1 = 0.0;
  if (e < len) {
    int b = q;
    for (; b + 8 < nparts; b += 16) {
      s0 += part[(size_t)b * len + e];
      s1 += part[(size_t)(b + 8) * len + e];
    }
    for (; b < nparts; b += 8) s0 += part[(size_t)b * len + e];
  }
  acc[q][o] = s0 + s1;
  __syncthreads();
  if (q == 0 && e < len) {
    double s = 0.0;
    for (int k = 0; k < 8; ++k) s += acc[k][o];
    out[e] = (float)((accumulate ? (double)out[e] : 0.0) + s * scale);
  }
}

// Data-gradient weights: Wf[t'][co][ci] = W[26 - t][ci][co] (reverse the taps,
// swap the channel roles).
__global__ void k_flip_weights(const uint16_t* __restrict__ W, uint16_t* __restrict__ Wf) {
  for (int e = threadIdx.x + blockIdx.x * blockDim.x; e < 27 * 256; e += blockDim.x * gridDim.x) {
    const int t = e / 256, ci = (e / 16) % 16, co = e % 16;
    Wf[((26 - t) * 16 + co) * 16 + ci] = W[(t * 16 + ci) * 16 + co];
  }
}

// Adam step on fp32 master weights; writes the bf16 copy used by the layers.
__global__ void k_adam(int n, float* __restrict__ w, const float* __restrict__ grad, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2,
                       const uint8_t* __restrict__ frozen, uint16_t* __restrict__ wbf) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (frozen && frozen[i]) {
      wbf[i] = f2bf(w[i]);
      continue;
    }
    const float gi = grad[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    wbf[i] = f2bf(w[i]);
  }
}

// MFMA lane-layout probe: D = A B for 16x16 bf16 A (row-major) and B.
__global__ void k_mfma_probe(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B, float* __restrict__ D) {
  const int lane = threadIdx.x & 63;
  v4s a, b;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    a[i] = (short)A[(lane % 16) * 16 + 4 * (lane / 16) + i];
    b[i] = (short)B[(4 * (lane / 16) + i) * 16 + lane % 16];
  }
  v4f d = {0.f, 0.f, 0.f, 0.f};
  d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) D[(4 * (lane / 16) + i) * 16 + lane % 16] = d[i];
}

// Tied layer weights from the two fp32 masters: (W0, W1, flip W1, flip W0),
// rounded to bfloat16.
__global__ void k_sym_weights(const float* __restrict__ w, uint16_t* __restrict__ layers) {
  for (int e = threadIdx.x + blockIdx.x * blockDim.x; e < 27 * 256; e += blockDim.x * gridDim.x) {
    const int t = e / 256, a = (e / 16) % 16, b = e % 16;
    const int ef = ((26 - t) * 16 + b) * 16 + a;  // flip(W)[t][a][b] = W[26 - t][b][a]
    layers[e] = f2bf(w[e]);
    layers[256 * 27 + e] = f2bf(w[256 * 27 + e]);
    layers[2 * 256 * 27 + e] = f2bf(w[256 * 27 + ef]);
    layers[3 * 256 * 27 + e] = f2bf(w[ef]);
  }
}

// Gradients of the two masters from the four layers' gradients:
// gW0 = g0 + flip g3, gW1 = g1 + flip g2 (flip is an involution).
__global__ void k_sym_grads(const float* __restrict__ gl, float* __restrict__ gw, int accumulate) {
  for (int e = threadIdx.x + blockIdx.x * blockDim.x; e < 27 * 256; e += blockDim.x * gridDim.x) {
    const int t = e / 256, a = (e / 16) % 16, b = e % 16;
    const int ef = ((26 - t) * 16 + b) * 16 + a;
    const float g0 = gl[e] + gl[3 * 256 * 27 + ef];
    const float g1 = gl[256 * 27 + e] + gl[2 * 256 * 27 + ef];
    gw[e] = accumulate ? gw[e] + g0 : g0;
    gw[256 * 27 + e] = accumulate ? gw[256 * 27 + e] + g1 : g1;
  }
}

// The preconditioner network: four masked linear layers (weights W0, W1,
// flip W1, flip W0 in the descriptor), the last one written in fp32.
// The four layers on nb grids at once (each act / out_f32 buffer holds nb
// grids of [voxel][16], one after the other).
int cnn_forward_layers(const cc_cnn_desc& net, uint16_t* act0, uint16_t* act1, uint16_t* act2, uint16_t* act3,
                       float* out_f32, hipStream_t s, int nb) {
  // The weights are staged with 16-byte loads.
  for (int l = 0; l < 4; ++l)
    if (((uintptr_t)net.weights[l] & 15) != 0) return CC_E_ARG;
  GridArgs g{net.nx, net.ny, net.nz, net.periodic[0], net.periodic[1], net.periodic[2]};
  const int tz = conv_tiles_z(net.nz);
  const int ipb = conv_ipb(net.nx, net.ny, net.nz, nb);
  const dim3 blocks = conv_grid(net.nx, net.ny, net.nz, nb, ipb);
  if (blocks.z > 65535 || blocks.y > 65535) return CC_E_ARG;
  const size_t bs = (size_t)net.nx * net.ny * net.nz * 16;
  uint16_t* acts[4] = {act0, act1, act2, act3};
  for (int l = 0; l < 3; ++l)
    hipLaunchKernelGGL(k_conv3<0>, blocks, dim3(kConvThreads), 0, s, g, acts[l], nullptr, net.weights[l], net.mask,
                       nullptr, acts[l + 1], nullptr, 0, tz, bs, nb, ipb);
  hipLaunchKernelGGL(k_conv3<0>, blocks, dim3(kConvThreads), 0, s, g, act3, nullptr, net.weights[3], net.mask,
                     nullptr, nullptr, out_f32, 0, tz, bs, nb, ipb);
  return (int)hipGetLastError();
}

// The same four layers from a single-channel float32 input grid u (channel
// 0 of the first layer's input; its other channels are zero) to channel 0 of
// the last layer as a single-channel float32 grid y, on nb grids of each.
int cnn_forward_ch0(const cc_cnn_desc& net, const float* u, uint16_t* act1, uint16_t* act2, uint16_t* act3, float* y,
                    hipStream_t s, int nb) {
  for (int l = 0; l < 4; ++l)
    if (((uintptr_t)net.weights[l] & 15) != 0) return CC_E_ARG;
  GridArgs g{net.nx, net.ny, net.nz, net.periodic[0], net.periodic[1], net.periodic[2]};
  const int tz = conv_tiles_z(net.nz);
  const int ipb = conv_ipb(net.nx, net.ny, net.nz, nb);
  const dim3 blocks = conv_grid(net.nx, net.ny, net.nz, nb, ipb);
  if (blocks.z > 65535 || blocks.y > 65535) return CC_E_ARG;
  const size_t bs = (size_t)net.nx * net.ny * net.nz * 16;
  hipLaunchKernelGGL(k_conv3_first, blocks, dim3(kConvThreads), 0, s, g, u, net.weights[0], net.mask, act1, tz, bs, nb,
                     ipb);
  hipLaunchKernelGGL(k_conv3<0>, blocks, dim3(kConvThreads), 0, s, g, act1, nullptr, net.weights[1], net.mask, nullptr,
                     act2, nullptr, 0, tz, bs, nb, ipb);
  hipLaunchKernelGGL(k_conv3<0>, blocks, dim3(kConvThreads), 0, s, g, act2, nullptr, net.weights[2], net.mask, nullptr,
                     act3, nullptr, 0, tz, bs, nb, ipb);
  hipLaunchKernelGGL(k_conv3<0>, blocks, dim3(kConvThreads), 0, s, g, act3, nullptr, net.weights[3], net.mask, nullptr,
                     nullptr, y, 1, tz, bs, nb, ipb);
  return (int)hipGetLastError();
}

}  // namespace cc

using namespace cc;

extern "C" {

int cc_mfma_probe(const uint16_t* A, const uint16_t* B, float* D, cc_stream_t stream) {
  if (!A || !B || !D) return CC_E_ARG;
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, A, B, D);
  return (int)hipGetLastError();
}

int cc_cnn_set_config(int32_t items_per_block) {
  if (items_per_block < 0 || items_per_block > 4096) return CC_E_ARG;
  g_conv_ipb = items_per_block;
  return 0;
}

int cc_cnn_conv3_forward(int32_t nx, int32_t ny, int32_t nz, const int32_t* periodic, const uint16_t* in,
                         const uint16_t* weights, const uint8_t* mask, const uint16_t* gate, int32_t relu,
                         uint16_t* out, float* out_f32, cc_stream_t stream) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !in || !weights || !mask || (!out && !out_f32)) return CC_E_ARG;
  // The weights are staged with 16-byte loads.
  if (((uintptr_t)weights & 15) != 0) return CC_E_ARG;
  GridArgs g{nx, ny, nz, periodic ? periodic[0] : 0, periodic ? periodic[1] : 0, periodic ? periodic[2] : 0};
  if ((nz + kCz - 1) / kCz > 65535 || (ny + kCy - 1) / kCy > 65535) return CC_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  const int tz = conv_tiles_z(nz);
  const int ipb = conv_ipb(nx, ny, nz, 1);
  const dim3 blocks = conv_grid(nx, ny, nz, 1, ipb);
  if (relu)
    hipLaunchKernelGGL(k_conv3<1>, blocks, dim3(kConvThreads), 0, s, g, in, nullptr, weights, mask, gate, out, out_f32,
                       0, tz, 0, 1, ipb);
  else
    hipLaunchKernelGGL(k_conv3<0>, blocks, dim3(kConvThreads), 0, s, g, in, nullptr, weights, mask, gate, out, out_f32,
                       0, tz, 0, 1, ipb);
  return (int)hipGetLastError();
}

int cc_cnn_flip_weights(const uint16_t* weights, uint16_t* flipped, cc_stream_t stream) {
  if (!weights || !flipped) return CC_E_ARG;
  hipLaunchKernelGGL(k_flip_weights, dim3(27), dim3(256), 0, (hipStream_t)stream, weights, flipped);
  return (int)hipGetLastError();
}

size_t cc_cnn_wgrad_workspace_bytes(int32_t nx, int32_t ny, int32_t nz, int32_t blocks) {
  (void)nx;
  (void)ny;
  (void)nz;
  return (size_t)(blocks > 0 ? blocks : 1) * 27 * 256 * sizeof(float);
}

int cc_cnn_conv3_wgrad(int32_t nx, int32_t ny, int32_t nz, const int32_t* periodic, const uint16_t* in,
                       const uint16_t* gout, float* dW, float scale, int32_t accumulate, int32_t blocks,
                       void* workspace, size_t workspace_bytes, cc_stream_t stream) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !in || !gout || !dW || !workspace || blocks <= 0) return CC_E_ARG;
  if (workspace_bytes < cc_cnn_wgrad_workspace_bytes(nx, ny, nz, blocks)) return CC_E_WORKSPACE;
  GridArgs g{nx, ny, nz, periodic ? periodic[0] : 0, periodic ? periodic[1] : 0, periodic ? periodic[2] : 0};
  const long long rows = (long long)((nx + 15) / 16) * ny * nz;
  const int rpb = (int)((rows + blocks - 1) / blocks);
  hipStream_t s = (hipStream_t)stream;
  float* part = (float*)workspace;
  hipLaunchKernelGGL(k_conv3_wgrad, dim3(blocks), dim3(256), 0, s, g, in, gout, part, rpb);
  hipLaunchKernelGGL(k_reduce_parts, dim3(27 * 256 / 32), dim3(256), 0, s, part, blocks, 27 * 256, dW, scale,
                     accumulate);
  return (int)hipGetLastError();
}

int cc_cnn_sym_weights(const float* w_master, uint16_t* layers, cc_stream_t stream) {
  if (!w_master || !layers) return CC_E_ARG;
  hipLaunchKernelGGL(k_sym_weights, dim3(27), dim3(256), 0, (hipStream_t)stream, w_master, layers);
  return (int)hipGetLastError();
}

int cc_cnn_sym_grads(const float* g_layers, float* g_master, int32_t accumulate, cc_stream_t stream) {
  if (!g_layers || !g_master) return CC_E_ARG;
  hipLaunchKernelGGL(k_sym_grads, dim3(27), dim3(256), 0, (hipStream_t)stream, g_layers, g_master, accumulate);
  return (int)hipGetLastError();
}

int cc_cnn_adam(int32_t n, float* w, const float* grad, float* m1, float* m2, float lr, float beta1, float beta2,
                float eps, int32_t step, const uint8_t* frozen, uint16_t* w_bf16, cc_stream_t stream) {
  if (n <= 0 || !w || !grad || !m1 || !m2 || !w_bf16 || step < 1) return CC_E_ARG;
  const float c1 = 1.f - powf(beta1, (float)step);
  const float c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, w, grad, m1, m2, lr, beta1,
                     beta2, eps, c1, c2, frozen, w_bf16);
  return (int)hipGetLastError();
}

}  // extern "C"
