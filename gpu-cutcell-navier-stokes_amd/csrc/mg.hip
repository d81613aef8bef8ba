// mg.hip - background-grid hierarchy of the multilevel learned preconditioner.
//
// Level l + 1 of a dof grid keeps every other grid point of level l: on a
// clamped axis n_c = (n_f + 1) / 2 and coarse point c sits on fine point 2c;
// on a periodic axis n_f is even and n_c = n_f / 2 (indices wrap). The
// prolongation P is tensor-product linear interpolation:
//   fine 2c      <- coarse c            (weight 1)
//   fine 2c + 1  <- coarse c, c + 1     (weights 1/2, 1/2; c + 1 wraps on a
//                                        periodic axis; past the end of a
//                                        clamped one, coarse c alone with
//                                        weight 1, so P keeps constants)
// and the restriction is its exact transpose R = P^T (coarse c gathers fine
// 2c - 1, 2c, 2c + 1 with weights 1/2, 1, 1/2 where the fine point exists,
// and 1 for the trailing fine point of an even-length clamped axis),
// so that  Q_l C^T C Q_l^T  is symmetric for any level operator C.
// Arrays are single-channel float32 grids (x fastest) read and written with
// an element stride, so a network's channel-0 output ([voxel][16]) can be
// prolongated in place.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {


// Restriction tap d in {-1, 0, 1} of coarse point c: the fine index and
// weight, or false when that fine point does not exist.
__device__ inline bool rest_tap(int c, int d, const MgAxis& ax, int* f, float* w) {
  int i = 2 * c + d;
  if (i < 0) {
    if (!ax.per) return false;
    i += ax.nf;
  } else if (i >= ax.nf) {
    if (!ax.per) return false;
    i -= ax.nf;
  }
  *f = i;
  // The last coarse point is the only source of a trailing odd fine point.
  *w = (d == 0 || (d == 1 && !ax.per && c + 1 == ax.nc)) ? 1.0f : 0.5f;
  return true;
}

// Batched over blockIdx.y: grid b of a batch starts b * fbs (fine) and
// b * cbs (coarse) elements in.
__global__ __launch_bounds__(256) void k_mg_restrict(MgDims g, const float* __restrict__ fine, int fs,
                                                     float* __restrict__ coarse, int cs, size_t fbs, size_t cbs) {
  fine += blockIdx.y * fbs;
  coarse += blockIdx.y * cbs;
  const int ncx = g.a[0].nc, ncy = g.a[1].nc, ncz = g.a[2].nc;
  const long long total = (long long)ncx * ncy * ncz;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < total;
       v += (long long)gridDim.x * blockDim.x) {
    const int cx = (int)(v % ncx), cy = (int)((v / ncx) % ncy), cz = (int)(v / ((long long)ncx * ncy));
    // Missing taps read the (always present) centre point at weight 0, so
    // every row's loads are independent and issued together; adding 0 * f
    // leaves the sum as the taps-only loop computes it.
    int fx[3], fy[3], fz[3];
    float wx[3], wy[3], wz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (!rest_tap(cx, d - 1, g.a[0], &fx[d], &wx[d])) {
        fx[d] = 2 * cx;
        wx[d] = 0.0f;
      }
      if (!rest_tap(cy, d - 1, g.a[1], &fy[d], &wy[d])) {
        fy[d] = 2 * cy;
        wy[d] = 0.0f;
      }
      if (!rest_tap(cz, d - 1, g.a[2], &fz[d], &wz[d])) {
        fz[d] = 2 * cz;
        wz[d] = 0.0f;
      }
    }
    float acc = 0.0f;
#pragma unroll
    for (int kz = 0; kz < 3; ++kz) {
#pragma unroll
      for (int ky = 0; ky < 3; ++ky) {
        const size_t rowoff = ((size_t)fz[kz] * g.a[1].nf + fy[ky]) * g.a[0].nf;
        float fv[3];
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) fv[kx] = fine[(rowoff + fx[kx]) * fs];
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) acc += (wz[kz] * wy[ky] * wx[kx]) * fv[kx];
      }
    }
    coarse[(size_t)v * cs] = acc;
  }
}

__global__ __launch_bounds__(256) void k_mg_prolong(MgDims g, const float* __restrict__ coarse, int cs,
                                                    const float* __restrict__ add, int as, float* __restrict__ fine,
                                                    size_t cbs, size_t abs_, size_t fbs) {
  coarse += blockIdx.y * cbs;
  if (add) add += blockIdx.y * abs_;
  fine += blockIdx.y * fbs;
  const int nfx = g.a[0].nf, nfy = g.a[1].nf, nfz = g.a[2].nf;
  const long long total = (long long)nfx * nfy * nfz;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < total;
       v += (long long)gridDim.x * blockDim.x) {
    const int fx = (int)(v % nfx), fy = (int)((v / nfx) % nfy), fz = (int)(v / ((long long)nfx * nfy));
    fine[v] = prolong_value(g, coarse, cs, add ? add[(size_t)v * as] : 0.0f, fx, fy, fz);
  }
}

__device__ inline uint16_t mg_f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__global__ __launch_bounds__(256) void k_mg_to_ch0(long long nvox, const float* __restrict__ src, int ss,
                                                   uint16_t* __restrict__ grid, size_t sbs, size_t gbs) {
  src += blockIdx.y * sbs;
  grid += blockIdx.y * gbs;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
       v += (long long)gridDim.x * blockDim.x)
    grid[(size_t)v * 16] = mg_f2bf(src[(size_t)v * ss]);
}

static int mg_grid(long long work) {
  long long g = (work + 255) / 256;
  if (g > 4096) g = 4096;
  return (int)(g < 1 ? 1 : g);
}

static bool mg_dims(const int32_t* nf, const int32_t* nc, const int32_t* per, MgDims* g) {
  if (!nf || !nc || !per) return false;
  for (int k = 0; k < 3; ++k) {
    if (nf[k] <= 0 || nc[k] <= 0) return false;
    if (per[k] ? (nf[k] % 2 != 0 || nc[k] != nf[k] / 2) : nc[k] != (nf[k] + 1) / 2) return false;
    g->a[k] = MgAxis{nf[k], nc[k], per[k] != 0};
  }
  return true;
}

int mg_restrict(const MgDims& g, const float* fine, int fs, float* coarse, int cs, hipStream_t s, int nb,
                size_t fbs, size_t cbs) {
  const long long nc = (long long)g.a[0].nc * g.a[1].nc * g.a[2].nc;
  hipLaunchKernelGGL(k_mg_restrict, dim3(mg_grid(nc), nb), dim3(256), 0, s, g, fine, fs, coarse, cs, fbs, cbs);
  return (int)hipGetLastError();
}

int mg_prolong(const MgDims& g, const float* coarse, int cs, const float* add, int as, float* fine, hipStream_t s,
               int nb, size_t cbs, size_t abs_, size_t fbs) {
  const long long nf = (long long)g.a[0].nf * g.a[1].nf * g.a[2].nf;
  hipLaunchKernelGGL(k_mg_prolong, dim3(mg_grid(nf), nb), dim3(256), 0, s, g, coarse, cs, add, as, fine, cbs, abs_,
                     fbs);
  return (int)hipGetLastError();
}

int mg_to_ch0(long long nvox, const float* src, int ss, uint16_t* grid, hipStream_t s, int nb, size_t sbs,
              size_t gbs) {
  hipLaunchKernelGGL(k_mg_to_ch0, dim3(mg_grid(nvox), nb), dim3(256), 0, s, nvox, src, ss, grid, sbs, gbs);
  return (int)hipGetLastError();
}

bool mg_level_dims(const cc_cnn_desc& fine, const cc_cnn_desc& coarse, MgDims* g) {
  const int32_t nf[3] = {fine.nx, fine.ny, fine.nz};
  const int32_t nc[3] = {coarse.nx, coarse.ny, coarse.nz};
  return mg_dims(nf, nc, fine.periodic, g);
}

}  // namespace cc

using namespace cc;

extern "C" {

int cc_mg_restrict(const int32_t* nf, const int32_t* nc, const int32_t* periodic, const float* fine,
                   int32_t fine_stride, float* coarse, int32_t coarse_stride, cc_stream_t stream) {
  MgDims g;
  if (!mg_dims(nf, nc, periodic, &g) || !fine || !coarse || fine_stride < 1 || coarse_stride < 1) return CC_E_ARG;
  return mg_restrict(g, fine, fine_stride, coarse, coarse_stride, (hipStream_t)stream, 1, 0, 0);
}

int cc_mg_prolong(const int32_t* nf, const int32_t* nc, const int32_t* periodic, const float* coarse,
                  int32_t coarse_stride, const float* add, int32_t add_stride, float* fine, cc_stream_t stream) {
  MgDims g;
  if (!mg_dims(nf, nc, periodic, &g) || !coarse || !fine || coarse_stride < 1 || (add && add_stride < 1))
    return CC_E_ARG;
  return mg_prolong(g, coarse, coarse_stride, add, add_stride, fine, (hipStream_t)stream, 1, 0, 0, 0);
}

int cc_mg_to_ch0(int64_t nvox, const float* src, int32_t src_stride, uint16_t* grid, cc_stream_t stream) {
  if (nvox <= 0 || !src || !grid || src_stride < 1) return CC_E_ARG;
  return mg_to_ch0(nvox, src, src_stride, grid, (hipStream_t)stream, 1, 0, 0);
}

}  // extern "C"
