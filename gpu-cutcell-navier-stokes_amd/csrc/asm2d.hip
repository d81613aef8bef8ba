// asm2d.hip - the 2D half-plane cut-cell mass and stiffness matrices H and K
// assembled on the device, in CSR.
//
// It follows the reference's assembly (code/oblique_cut_sbp.py:271-433): the
// unit square is N x N cells of tensor B-splines of degree p on open uniform
// knots, cut by the half plane n . x >= offset. A cell whose four corners are
// inside is full and contributes the tensor products of the 1D element forms
// (H_e = My x Mx, K_e = My x Kx + Ky x Mx); a cell with some corner inside is
// cut: its square is clipped against the half plane (Sutherland-Hodgman,
// inside tolerance 1e-14, crossings clamped to the edge, duplicate vertices
// closer than 1e-13 dropped), fanned into triangles from the first vertex,
// each integrated by the q x q collapsed (Duffy) product Gauss rule, and the
// local Gram blocks H[a][b] = sum w phi_a phi_b, K[a][b] = sum w grad phi_a .
// grad phi_b over its (p+1)^2 local functions (a = ay (p+1) + ax) are summed
// at the points; a clipped polygon with fewer than 3 vertices or area at most
// 1e-16 contributes nothing. The dof of basis (ix, iy) is iy * nb + ix.
//
// Kernels:
//   k_cut2d_gram     one workgroup per cut cell (cells of many systems in one
//                    launch): clip, rule, basis values and first derivatives
//                    (Cox-de Boor on the knot span), the two Gram blocks;
//   k_asm2d_count    per row, the stencil offsets (|dx|, |dy| <= p) some
//                    contributing cell covers (the host's pattern);
//   k_asm2d_scan     one block: row pointers;
//   k_asm2d_fill     per row and offset: the column and the H and K values,
//                    each summed over the covering cells in a fixed order
//                    (ey, then ex, ascending), so no atomics are used.
// Products that decide inside tests are formed without fused multiply-adds,
// as the host forms them.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/cutcell_hip.h"
#include "common.h"

namespace cc {

namespace {

constexpr int kMaxP1 = 6;          // p <= 5
constexpr int kMaxPoly = 8;        // a clipped square has at most 5 vertices
constexpr int kGramThreads = 256;
constexpr int kPtsChunk = 256;

__device__ inline double dot2(double n0, double n1, double x, double y) {
  return __dadd_rn(__dmul_rn(n0, x), __dmul_rn(n1, y));
}

__device__ inline double knot(int k, int N, int p) {
  // Open uniform knots: p + 1 zeros, k h for k = 1 .. N - 1, p + 1 ones.
  const int j = k - p;
  if (j <= 0) return 0.0;
  if (j >= N) return 1.0;
  return (double)j * (1.0 / (double)N);
}

// Values and first derivatives of the p + 1 basis functions e .. e + p that
// are nonzero on element e (knot span s = e + p), at x (Cox-de Boor).
__device__ inline void basis_on_element(int e, int N, int p, double x, double* val, double* der) {
  const int s = e + p;
  double left[kMaxP1], right[kMaxP1], nd[kMaxP1];
  nd[0] = 1.0;
  double prev[kMaxP1];
  for (int j = 1; j <= p; ++j) {
    left[j] = x - knot(s + 1 - j, N, p);
    right[j] = knot(s + j, N, p) - x;
    double saved = 0.0;
    if (j == p)
      for (int r = 0; r < p; ++r) prev[r] = nd[r];  // degree p - 1 values
    for (int r = 0; r < j; ++r) {
      const double t = nd[r] / (right[r + 1] + left[j - r]);
      nd[r] = saved + right[r + 1] * t;
      saved = left[j - r] * t;
    }
    nd[j] = saved;
  }
  for (int r = 0; r <= p; ++r) val[r] = nd[r];
  if (p == 0) {
    der[0] = 0.0;
    return;
  }
  // N'_{i,p} = p / (t_{i+p} - t_i) N_{i,p-1} - p / (t_{i+p+1} - t_{i+1}) N_{i+1,p-1}
  // for i = e .. e + p; the degree-(p-1) functions on the span are e+1 .. e+p.
  for (int r = 0; r <= p; ++r) {
    const int i = e + r;
    double d = 0.0;
    if (r >= 1) {
      const double den = knot(i + p, N, p) - knot(i, N, p);
      if (den > 0.0) d += (double)p / den * prev[r - 1];
    }
    if (r <= p - 1) {
      const double den = knot(i + p + 1, N, p) - knot(i + 1, N, p);
      if (den > 0.0) d -= (double)p / den * prev[r];
    }
    der[r] = d;
  }
}

struct Clip {
  double x[kMaxPoly], y[kMaxPoly];
  int n;
};

__device__ inline bool inside(double n0, double n1, double off, double x, double y) {
  return dot2(n0, n1, x, y) >= off - 1.0e-14;
}

__device__ inline void crossing(double n0, double n1, double off, double ax, double ay, double bx, double by,
                                double* cx, double* cy) {
  const double fa = dot2(n0, n1, ax, ay) - off;
  const double fb = dot2(n0, n1, bx, by) - off;
  const double denom = fa - fb;
  if (fabs(denom) < 1.0e-30) {
    *cx = 0.5 * (ax + bx);
    *cy = 0.5 * (ay + by);
    return;
  }
  double t = fa / denom;
  t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  *cx = __dadd_rn(ax, __dmul_rn(t, bx - ax));
  *cy = __dadd_rn(ay, __dmul_rn(t, by - ay));
}

__device__ inline void clip_square(double x0, double x1, double y0, double y1, double n0, double n1, double off,
                                   Clip& out) {
  const double px[4] = {x0, x1, x1, x0}, py[4] = {y0, y0, y1, y1};
  double ox[kMaxPoly], oy[kMaxPoly];
  int m = 0;
  double prx = px[3], pry = py[3];
  bool prev_in = inside(n0, n1, off, prx, pry);
  for (int k = 0; k < 4; ++k) {
    const double cx = px[k], cy = py[k];
    const bool cur_in = inside(n0, n1, off, cx, cy);
    if (cur_in) {
      if (!prev_in) {
        crossing(n0, n1, off, prx, pry, cx, cy, &ox[m], &oy[m]);
        ++m;
      }
      ox[m] = cx;
      oy[m] = cy;
      ++m;
    } else if (prev_in) {
      crossing(n0, n1, off, prx, pry, cx, cy, &ox[m], &oy[m]);
      ++m;
    }
    prx = cx;
    pry = cy;
    prev_in = cur_in;
  }
  // Drop vertices closer than 1e-13 to the last kept one, then a closing
  // duplicate.
  int n = 0;
  for (int k = 0; k < m; ++k) {
    if (n > 0) {
      const double dx = ox[k] - out.x[n - 1], dy = oy[k] - out.y[n - 1];
      if (sqrt(dx * dx + dy * dy) <= 1.0e-13) continue;
    }
    out.x[n] = ox[k];
    out.y[n] = oy[k];
    ++n;
  }
  if (n > 1) {
    const double dx = out.x[0] - out.x[n - 1], dy = out.y[0] - out.y[n - 1];
    if (sqrt(dx * dx + dy * dy) <= 1.0e-13) --n;
  }
  out.n = n;
}

__device__ inline double poly_area(const Clip& c) {
  if (c.n < 3) return 0.0;
  double s = 0.0;
  for (int k = 0; k < c.n; ++k) {
    const int k1 = (k + 1) % c.n;
    s += c.x[k] * c.y[k1] - c.y[k] * c.x[k1];
  }
  return 0.5 * s;
}

__global__ __launch_bounds__(kGramThreads) void k_cut2d_gram(int ncells, const cc_cut2d_cell* __restrict__ cells,
                                                             int N, int p, int q, const double* __restrict__ gx,
                                                             const double* __restrict__ gw, double* __restrict__ Hb,
                                                             double* __restrict__ Kb, double* __restrict__ area_out) {
  const int c = blockIdx.x;
  if (c >= ncells) return;
  const int p1 = p + 1;
  const int nloc = p1 * p1;
  const int npair = nloc * nloc;
  __shared__ Clip poly;
  __shared__ double s_area;
  __shared__ double pv[kPtsChunk][4][kMaxP1];  // vx, dvx, vy, dvy per point
  __shared__ double pw[kPtsChunk];
  const cc_cut2d_cell cell = cells[c];
  const double h = 1.0 / (double)N;
  const double x0 = cell.ix * h, x1 = (cell.ix + 1) * h, y0 = cell.iy * h, y1 = (cell.iy + 1) * h;
  if (threadIdx.x == 0) {
    Clip cl;
    clip_square(x0, x1, y0, y1, cell.normal_x, cell.normal_y, cell.offset, cl);
    double a = poly_area(cl);
    if (a < 0.0) {  // orient counter-clockwise
      for (int k = 0; k < cl.n / 2; ++k) {
        const double tx = cl.x[k], ty = cl.y[k];
        cl.x[k] = cl.x[cl.n - 1 - k];
        cl.y[k] = cl.y[cl.n - 1 - k];
        cl.x[cl.n - 1 - k] = tx;
        cl.y[cl.n - 1 - k] = ty;
      }
    }
    poly = cl;
    s_area = (cl.n < 3 || fabs(a) <= 1.0e-16) ? 0.0 : fabs(a);
  }
  __syncthreads();
  // Up to 6 (a, b) pairs per thread (p <= 5: 1296 pairs, 256 threads).
  constexpr int kPairs = (kMaxP1 * kMaxP1 * kMaxP1 * kMaxP1 + kGramThreads - 1) / kGramThreads;
  double accH[kPairs], accK[kPairs];
#pragma unroll
  for (int k = 0; k < kPairs; ++k) accH[k] = accK[k] = 0.0;
  const bool skip = s_area == 0.0;
  const int ntri = skip ? 0 : poly.n - 2;
  const int per_tri = q * q;
  const int npts = ntri * per_tri;
  for (int base = 0; base < npts; base += kPtsChunk) {
    const int pt = base + threadIdx.x;
    if (pt < npts) {
      const int tri = pt / per_tri, r = pt % per_tri;
      const int i = r / q, j = r % q;  // U = u_i (outer), V = u_j (inner)
      const double v0x = poly.x[0], v0y = poly.y[0];
      const double e1x = poly.x[tri + 1] - v0x, e1y = poly.y[tri + 1] - v0y;
      const double e2x = poly.x[tri + 2] - v0x, e2y = poly.y[tri + 2] - v0y;
      const double det = fabs(e1x * e2y - e1y * e2x);
      const double U = 0.5 * (gx[i] + 1.0), V = 0.5 * (gx[j] + 1.0);
      const double W = det * (1.0 - U) * (0.5 * gw[i]) * (0.5 * gw[j]);
      const double xq = v0x + U * e1x + ((1.0 - U) * V) * e2x;
      const double yq = v0y + U * e1y + ((1.0 - U) * V) * e2y;
      basis_on_element(cell.ix, N, p, xq, pv[threadIdx.x][0], pv[threadIdx.x][1]);
      basis_on_element(cell.iy, N, p, yq, pv[threadIdx.x][2], pv[threadIdx.x][3]);
      pw[threadIdx.x] = W;
    }
    __syncthreads();
    const int cnt = min(kPtsChunk, npts - base);
#pragma unroll
    for (int k = 0; k < kPairs; ++k) {
      const int pr = threadIdx.x + k * kGramThreads;
      if (pr < npair) {
        const int a = pr / nloc, b = pr % nloc;
        const int ay = a / p1, ax = a % p1, by = b / p1, bx = b % p1;
        double h_ = 0.0, k_ = 0.0;
        for (int t = 0; t < cnt; ++t) {
          const double w = pw[t];
          const double vxa = pv[t][0][ax], dxa = pv[t][1][ax], vya = pv[t][2][ay], dya = pv[t][3][ay];
          const double vxb = pv[t][0][bx], dxb = pv[t][1][bx], vyb = pv[t][2][by], dyb = pv[t][3][by];
          const double pa = vya * vxa, pb = vyb * vxb;
          h_ += (pa * w) * pb;
          k_ += ((vya * dxa) * w) * (vyb * dxb) + ((dya * vxa) * w) * (dyb * vxb);
        }
        accH[k] += h_;
        accK[k] += k_;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < kPairs; ++k) {
    const int pr = threadIdx.x + k * kGramThreads;
    if (pr < npair) {
      Hb[(size_t)c * npair + pr] = accH[k];
      Kb[(size_t)c * npair + pr] = accK[k];
    }
  }
  if (threadIdx.x == 0) area_out[c] = s_area;
}

// Cell class: 0 empty, 1 full, 2 cut with a contributing polygon, 3 cut but
// degenerate (contributes nothing).
__device__ inline bool contributes(int cls) { return cls == 1 || cls == 2; }

__global__ __launch_bounds__(256) void k_asm2d_count(int N, int p, const int8_t* __restrict__ cls,
                                                     int* __restrict__ rowcnt) {
  const int nb = N + p;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb * nb) return;
  const int ix = i % nb, iy = i / nb;
  int cnt = 0;
  for (int dy = -p; dy <= p; ++dy) {
    const int jy = iy + dy;
    if (jy < 0 || jy >= nb) continue;
    const int ey0 = max(max(iy, jy) - p, 0), ey1 = min(min(iy, jy), N - 1);
    for (int dx = -p; dx <= p; ++dx) {
      const int jx = ix + dx;
      if (jx < 0 || jx >= nb) continue;
      const int ex0 = max(max(ix, jx) - p, 0), ex1 = min(min(ix, jx), N - 1);
      bool any = false;
      for (int ey = ey0; ey <= ey1 && !any; ++ey)
        for (int ex = ex0; ex <= ex1; ++ex)
          if (contributes(cls[ey * N + ex])) {
            any = true;
            break;
          }
      cnt += any;
    }
  }
  rowcnt[i] = cnt;
}

// indptr[0] = 0, indptr[i + 1] = indptr[i] + rowcnt[i] (one block).
__global__ __launch_bounds__(1024) void k_asm2d_scan(int n, const int* __restrict__ rowcnt, int* __restrict__ indptr) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int lo = min(n, t * per), hi = min(n, lo + per);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += rowcnt[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t == 0 ? 0 : part[t - 1];
  if (t == 0) indptr[0] = 0;
  for (int i = lo; i < hi; ++i) {
    run += rowcnt[i];
    indptr[i + 1] = run;
  }
}

__global__ __launch_bounds__(256) void k_asm2d_fill(int N, int p, const int8_t* __restrict__ cls,
                                                    const int* __restrict__ slot, const double* __restrict__ Me,
                                                    const double* __restrict__ Ke, const double* __restrict__ Hb,
                                                    const double* __restrict__ Kb, const int* __restrict__ indptr,
                                                    int* __restrict__ indices, double* __restrict__ Hv,
                                                    double* __restrict__ Kv) {
  const int nb = N + p, p1 = p + 1, nloc = p1 * p1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb * nb) return;
  const int ix = i % nb, iy = i / nb;
  int pos = indptr[i];
  auto M1 = [&](int e, int a, int b) { return Me[((size_t)e * p1 + a) * p1 + b]; };
  auto K1 = [&](int e, int a, int b) { return Ke[((size_t)e * p1 + a) * p1 + b]; };
  for (int dy = -p; dy <= p; ++dy) {
    const int jy = iy + dy;
    if (jy < 0 || jy >= nb) continue;
    const int ey0 = max(max(iy, jy) - p, 0), ey1 = min(min(iy, jy), N - 1);
    for (int dx = -p; dx <= p; ++dx) {
      const int jx = ix + dx;
      if (jx < 0 || jx >= nb) continue;
      const int ex0 = max(max(ix, jx) - p, 0), ex1 = min(min(ix, jx), N - 1);
      bool any = false;
      double hs = 0.0, ks = 0.0;
      for (int ey = ey0; ey <= ey1; ++ey) {
        const int ay = iy - ey, by = jy - ey;
        for (int ex = ex0; ex <= ex1; ++ex) {
          const int c = cls[ey * N + ex];
          if (!contributes(c)) continue;
          any = true;
          const int ax = ix - ex, bx = jx - ex;
          if (c == 1) {
            const double my = M1(ey, ay, by), mx = M1(ex, ax, bx);
            hs += my * mx;
            ks += my * K1(ex, ax, bx) + K1(ey, ay, by) * mx;
          } else {
            const size_t o = (size_t)slot[ey * N + ex] * nloc * nloc + (size_t)(ay * p1 + ax) * nloc + (by * p1 + bx);
            hs += Hb[o];
            ks += Kb[o];
          }
        }
      }
      if (any) {
        indices[pos] = jy * nb + jx;
        Hv[pos] = hs;
        Kv[pos] = ks;
        ++pos;
      }
    }
  }
}

}  // namespace

}  // namespace cc

using namespace cc;

extern "C" {

int cc_cut2d_gram(int32_t ncells, const cc_cut2d_cell* cells, int32_t n_elements, int32_t degree, int32_t q,
                  const double* gauss_x, const double* gauss_w, double* H_blocks, double* K_blocks, double* areas,
                  cc_stream_t stream) {
  if (ncells < 0 || n_elements <= 0 || degree < 1 || degree > kMaxP1 - 1 || q <= 0) return CC_E_ARG;
  if (ncells == 0) return 0;
  if (!cells || !gauss_x || !gauss_w || !H_blocks || !K_blocks || !areas) return CC_E_ARG;
  hipLaunchKernelGGL(k_cut2d_gram, dim3(ncells), dim3(kGramThreads), 0, (hipStream_t)stream, ncells, cells,
                     n_elements, degree, q, gauss_x, gauss_w, H_blocks, K_blocks, areas);
  return (int)hipGetLastError();
}

int cc_asm2d_count(int32_t n_elements, int32_t degree, const int8_t* cell_class, int32_t* row_counts,
                   int32_t* indptr, cc_stream_t stream) {
  if (n_elements <= 0 || degree < 1 || degree > kMaxP1 - 1 || !cell_class || !row_counts || !indptr) return CC_E_ARG;
  const int nb = n_elements + degree, n = nb * nb;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_asm2d_count, dim3((n + 255) / 256), dim3(256), 0, s, n_elements, degree, cell_class,
                     row_counts);
  hipLaunchKernelGGL(k_asm2d_scan, dim3(1), dim3(1024), 0, s, n, row_counts, indptr);
  return (int)hipGetLastError();
}

int cc_asm2d_fill(int32_t n_elements, int32_t degree, const int8_t* cell_class, const int32_t* cut_slot,
                  const double* elem_mass, const double* elem_stiff, const double* H_blocks, const double* K_blocks,
                  const int32_t* indptr, int32_t* indices, double* H_vals, double* K_vals, cc_stream_t stream) {
  if (n_elements <= 0 || degree < 1 || degree > kMaxP1 - 1 || !cell_class || !cut_slot || !elem_mass ||
      !elem_stiff || !indptr || !indices || !H_vals || !K_vals)
    return CC_E_ARG;
  const int nb = n_elements + degree, n = nb * nb;
  hipLaunchKernelGGL(k_asm2d_fill, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n_elements, degree,
                     cell_class, cut_slot, elem_mass, elem_stiff, H_blocks, K_blocks, indptr, indices, H_vals,
                     K_vals);
  return (int)hipGetLastError();
}

}  // extern "C"
