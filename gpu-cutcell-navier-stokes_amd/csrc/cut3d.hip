// cut3d.hip - local Gram blocks of 3D level-set cut cells on the device.
//
// Per cut cell it computes what the host's levelset_cut_data does
// (cutcell/grid3d.py): the cell's six Kuhn tetrahedra (vertex 0, then one
// axis bit added at a time in each of the six axis orders, then the far
// corner) are clipped to the fluid side phi >= 0 of the level set's linear
// interpolant from the eight corner values, each clipped piece split into
// tetrahedra exactly as the host splits it:
//   one vertex inside:    (P_i, X_i,o0, X_i,o1, X_i,o2)
//   two inside (i0 < i1): the prism (P_i0, X_i0,o0, X_i0,o1 | P_i1, X_i1,o0, X_i1,o1)
//   three inside:         the prism (P_i0, P_i1, P_i2 | X_i0,o, X_i1,o, X_i2,o)
//   prism (a0 a1 a2 | b0 b1 b2) -> (a0 a1 a2 b2), (a0 a1 b1 b2), (a0 b0 b1 b2)
// (X_a,b the crossing on edge a-b, P_a + t (P_b - P_a) with t = f_a / (f_a - f_b);
// outside vertices o ascending). A sub-tetrahedron (v0, v1, v2, v3) maps the
// collapsed Gauss rule as v0 + xi1 e1 + xi2 e2 + xi3 e3 (e_k = v_k - v0) with
// weight |det (e1, e2, e3)| w. The quadrature is not exact for these
// integrands, so the sub-tetrahedra and their vertex orders must be the
// host's; only the order of the sums differs.
//
// At each point the p + 1 B-splines alive on the cell's element along each
// axis are evaluated by Cox-de Boor on the axis's knots (the point clamped to
// the element), with first derivatives, and
//   mass[i][j]  += w (phi_i phi_j)
//   stiff[i][j] += w (gx_i gx_j + gy_i gy_j + gz_i gz_j)
// for the (p + 1)^3 local functions i = (iz, iy, ix) (x fastest). Each
// product is commutative, so the blocks come out exactly symmetric.
// One workgroup per cut cell; 256 points at a time are staged in LDS, each
// thread accumulating its share of the block entries over them.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {
namespace {

constexpr int kP1 = 4;    // degree <= 3
constexpr int kSub = 18;  // six tetrahedra, at most three pieces each
constexpr int kNT = 256;  // threads per cut cell = points per stage
constexpr int kEnt = (kP1 * kP1 * kP1 * kP1 * kP1 * kP1 + kNT - 1) / kNT;  // 16 entries per thread

struct V3 {
  double x, y, z;
};

__device__ inline V3 crossing(const V3& a, const V3& b, double fa, double fb) {
  const double t = fa / (fa - fb);
  return V3{a.x + t * (b.x - a.x), a.y + t * (b.y - a.y), a.z + t * (b.z - a.z)};
}

// Values and d/dt of the p + 1 basis functions alive on knot span `span` at
// parameter t (Cox-de Boor, the host's rounding order).
__device__ inline void basis_span(const double* __restrict__ kn, int span, int p, double t, double* val,
                                  double* der) {
  double nd[kP1], prev[kP1], left[kP1], right[kP1];
  nd[0] = 1.0;
  prev[0] = 1.0;
  for (int j = 1; j <= p; ++j) {
    left[j] = t - kn[span + 1 - j];
    right[j] = kn[span + j] - t;
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      const double den = right[r + 1] + left[j - r];
      const double tmp = den != 0.0 ? nd[r] / den : 0.0;
      nd[r] = saved + right[r + 1] * tmp;
      saved = left[j - r] * tmp;
    }
    nd[j] = saved;
    if (j == p - 1)
      for (int r = 0; r <= j; ++r) prev[r] = nd[r];  // the degree p - 1 values
  }
  for (int a = 0; a <= p; ++a) {
    val[a] = nd[a];
    const int i = span - p + a;
    double d = 0.0;
    if (p >= 1 && a >= 1) {
      const double den = kn[i + p] - kn[i];
      if (den > 0.0) d += (double)p * prev[a - 1] / den;
    }
    if (p >= 1 && a <= p - 1) {
      const double den = kn[i + p + 1] - kn[i + 1];
      if (den > 0.0) d -= (double)p * prev[a] / den;
    }
    der[a] = d;
  }
}

__global__ __launch_bounds__(kNT) void k_cut3d_gram(const cc_cut3d_desc d, double* __restrict__ mass,
                                                    double* __restrict__ stiff, double* __restrict__ frac) {
  __shared__ V3 sv[kSub][4];
  __shared__ double sdet[kSub];
  __shared__ int scount[6];
  __shared__ int slist[kSub];
  __shared__ int snsub;
  __shared__ double pv[3][kP1][kNT];
  __shared__ double pd[3][kP1][kNT];
  __shared__ double pw[kNT];

  const int s = blockIdx.x;
  const int tid = threadIdx.x;
  const int p = d.degree, p1 = p + 1, P = p1 * p1 * p1;
  const int ec[3] = {d.cells[3 * s], d.cells[3 * s + 1], d.cells[3 * s + 2]};
  double h[3];
  for (int k = 0; k < 3; ++k) h[k] = d.length[k] / (double)d.ne[k];
  const V3 base{(double)ec[0] * h[0], (double)ec[1] * h[1], (double)ec[2] * h[2]};

  if (tid < 6) {
    // Kuhn tetrahedron tid: axis order (tid's permutation of x, y, z).
    const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int cv[4];
    cv[0] = 0;
    int acc = 0;
    for (int m = 0; m < 3; ++m) {
      acc |= 1 << perm[tid][m];
      cv[m + 1] = acc;
    }
    V3 P4[4];
    double F[4];
    for (int v = 0; v < 4; ++v) {
      const int c = cv[v];
      P4[v] = V3{base.x + h[0] * (double)(c & 1), base.y + h[1] * (double)((c >> 1) & 1),
                 base.z + h[2] * (double)((c >> 2) & 1)};
      F[v] = d.corner[8 * (size_t)s + c];
    }
    int ins[4], outs[4], ni = 0, no = 0;
    for (int v = 0; v < 4; ++v) {
      if (F[v] >= 0.0)
        ins[ni++] = v;
      else
        outs[no++] = v;
    }
    V3(*o)[4] = &sv[3 * tid];
    int cnt = 0;
    auto prism = [&](V3 a0, V3 a1, V3 a2, V3 b0, V3 b1, V3 b2) {
      o[0][0] = a0; o[0][1] = a1; o[0][2] = a2; o[0][3] = b2;
      o[1][0] = a0; o[1][1] = a1; o[1][2] = b1; o[1][3] = b2;
      o[2][0] = a0; o[2][1] = b0; o[2][2] = b1; o[2][3] = b2;
      cnt = 3;
    };
    if (ni == 4) {
      for (int v = 0; v < 4; ++v) o[0][v] = P4[v];
      cnt = 1;
    } else if (ni == 1) {
      const int i = ins[0];
      o[0][0] = P4[i];
      for (int m = 0; m < 3; ++m) o[0][m + 1] = crossing(P4[i], P4[outs[m]], F[i], F[outs[m]]);
      cnt = 1;
    } else if (ni == 3) {
      const int q = outs[0];
      prism(P4[ins[0]], P4[ins[1]], P4[ins[2]], crossing(P4[ins[0]], P4[q], F[ins[0]], F[q]),
            crossing(P4[ins[1]], P4[q], F[ins[1]], F[q]), crossing(P4[ins[2]], P4[q], F[ins[2]], F[q]));
    } else if (ni == 2) {
      const int i0 = ins[0], i1 = ins[1], o0 = outs[0], o1 = outs[1];
      prism(P4[i0], crossing(P4[i0], P4[o0], F[i0], F[o0]), crossing(P4[i0], P4[o1], F[i0], F[o1]), P4[i1],
            crossing(P4[i1], P4[o0], F[i1], F[o0]), crossing(P4[i1], P4[o1], F[i1], F[o1]));
    }
    scount[tid] = cnt;
  }
  __syncthreads();
  if (tid == 0) {
    int n = 0;
    for (int t = 0; t < 6; ++t)
      for (int m = 0; m < scount[t]; ++m) slist[n++] = 3 * t + m;
    snsub = n;
  }
  __syncthreads();
  const int nsub = snsub;
  if (tid < nsub) {
    const V3* v = sv[slist[tid]];
    const double e1x = v[1].x - v[0].x, e1y = v[1].y - v[0].y, e1z = v[1].z - v[0].z;
    const double e2x = v[2].x - v[0].x, e2y = v[2].y - v[0].y, e2z = v[2].z - v[0].z;
    const double e3x = v[3].x - v[0].x, e3y = v[3].y - v[0].y, e3z = v[3].z - v[0].z;
    const double cx = e2y * e3z - e2z * e3y, cy = e2z * e3x - e2x * e3z, cz = e2x * e3y - e2y * e3x;
    sdet[tid] = fabs(e1x * cx + e1y * cy + e1z * cz);
  }
  __syncthreads();
  if (tid == 0) {
    double vol = 0.0;
    for (int i = 0; i < nsub; ++i) vol += sdet[i] / 6.0;
    frac[s] = vol / (h[0] * h[1] * h[2]);
  }

  double am[kEnt], ak[kEnt];
  for (int e = 0; e < kEnt; ++e) am[e] = ak[e] = 0.0;
  const int nq = d.nq;
  const int npts = nsub * nq;
  for (int pt0 = 0; pt0 < npts; pt0 += kNT) {
    __syncthreads();  // the previous stage's values are consumed
    const int pt = pt0 + tid;
    if (pt < npts) {
      const int sl = pt / nq, qi = pt % nq;
      const V3* v = sv[slist[sl]];
      const double r0 = d.rule_pts[3 * qi], r1 = d.rule_pts[3 * qi + 1], r2 = d.rule_pts[3 * qi + 2];
      const double x[3] = {v[0].x + r0 * (v[1].x - v[0].x) + r1 * (v[2].x - v[0].x) + r2 * (v[3].x - v[0].x),
                           v[0].y + r0 * (v[1].y - v[0].y) + r1 * (v[2].y - v[0].y) + r2 * (v[3].y - v[0].y),
                           v[0].z + r0 * (v[1].z - v[0].z) + r1 * (v[2].z - v[0].z) + r2 * (v[3].z - v[0].z)};
      pw[tid] = sdet[sl] * d.rule_w[qi];
      for (int k = 0; k < 3; ++k) {
        const double lo = (double)ec[k] * h[k], hi = (double)(ec[k] + 1) * h[k];
        const double xq = fmin(fmax(x[k], lo), hi);
        double val[kP1], der[kP1];
        basis_span(d.knots[k], ec[k] + p, p, xq / d.length[k], val, der);
        for (int a = 0; a < p1; ++a) {
          pv[k][a][tid] = val[a];
          pd[k][a][tid] = der[a] / d.length[k];
        }
      }
    } else {
      pw[tid] = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int a = 0; a < p1; ++a) pv[k][a][tid] = pd[k][a][tid] = 0.0;
    }
    __syncthreads();
    const int nb = min(kNT, npts - pt0);
    // Entry by entry over the staged points: one entry's indices live at a
    // time, the accumulators of all of them across the stages.
#pragma unroll
    for (int e = 0; e < kEnt; ++e) {
      const int ent = tid + kNT * e;
      if (ent >= P * P) break;
      const int i = ent / P, j = ent % P;
      const int iz = i / (p1 * p1), iy = (i / p1) % p1, ix = i % p1;
      const int jz = j / (p1 * p1), jy = (j / p1) % p1, jx = j % p1;
      double m = am[e], kk = ak[e];
#pragma unroll 2
      for (int b = 0; b < nb; ++b) {
        const double w = pw[b];
        const double vzi = pv[2][iz][b], vyi = pv[1][iy][b], vxi = pv[0][ix][b];
        const double vzj = pv[2][jz][b], vyj = pv[1][jy][b], vxj = pv[0][jx][b];
        const double dzi = pd[2][iz][b], dyi = pd[1][iy][b], dxi = pd[0][ix][b];
        const double dzj = pd[2][jz][b], dyj = pd[1][jy][b], dxj = pd[0][jx][b];
        const double phi_i = __dmul_rn(__dmul_rn(vzi, vyi), vxi), phi_j = __dmul_rn(__dmul_rn(vzj, vyj), vxj);
        const double gxi = __dmul_rn(__dmul_rn(vzi, vyi), dxi), gxj = __dmul_rn(__dmul_rn(vzj, vyj), dxj);
        const double gyi = __dmul_rn(__dmul_rn(vzi, dyi), vxi), gyj = __dmul_rn(__dmul_rn(vzj, dyj), vxj);
        const double gzi = __dmul_rn(__dmul_rn(dzi, vyi), vxi), gzj = __dmul_rn(__dmul_rn(dzj, vyj), vxj);
        m = __fma_rn(w, __dmul_rn(phi_i, phi_j), m);
        const double gg = __dadd_rn(__dadd_rn(__dmul_rn(gxi, gxj), __dmul_rn(gyi, gyj)), __dmul_rn(gzi, gzj));
        kk = __fma_rn(w, gg, kk);
      }
      am[e] = m;
      ak[e] = kk;
    }
  }
#pragma unroll
  for (int e = 0; e < kEnt; ++e) {
    const int ent = tid + kNT * e;
    if (ent < P * P) {
      mass[(size_t)s * P * P + ent] = am[e];
      stiff[(size_t)s * P * P + ent] = ak[e];
    }
  }
}

}  // namespace
}  // namespace cc

using namespace cc;

extern "C" {

int cc_cut3d_gram(const cc_cut3d_desc* d, double* mass, double* stiff, double* frac, cc_stream_t stream) {
  if (!d || d->n_cut < 0 || d->degree < 1 || d->degree > kP1 - 1 || !mass || !stiff || !frac) return CC_E_ARG;
  if (d->n_cut == 0) return 0;
  if (!d->cells || !d->corner || !d->rule_pts || !d->rule_w || d->nq < 1) return CC_E_ARG;
  for (int k = 0; k < 3; ++k) {
    if (!d->knots[k] || d->ne[k] < 1 || !(d->length[k] > 0.0)) return CC_E_ARG;
    // The span of the last element reads knots up to index ne + 2 p.
    if (d->n_knots[k] < d->ne[k] + 2 * d->degree + 1) return CC_E_ARG;
  }
  hipLaunchKernelGGL(k_cut3d_gram, dim3((unsigned)d->n_cut), dim3(kNT), 0, (hipStream_t)stream, *d, mass, stiff, frac);
  return (int)hipGetLastError();
}

}  // extern "C"
