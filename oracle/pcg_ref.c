/*
 * pcg_ref.c - CPU oracle for the cut-cell PCG path (test infrastructure only).
 *
 * This file is the checker, never the product: only the test suite, the smoke
 * check and the CPU-baseline leg of the benchmark load it. It restates, in
 * plain scalar C with float64 arithmetic and a fixed left-to-right summation
 * order, the two Jacobi-PCG loops of the reference:
 *
 *  - oracle_pcg_gpu_loop: the reference GPU loop, solve_gpu_cg,
 *    code/fast_solver_backends.py:233-274 (reference repository):
 *      x = 0; r = b; d = diag(A); z = (|d| > 0) ? r / d : r; p = z; rho = r.z;
 *      bnorm = ||b|| (0 -> 1);
 *      for k = 1..maxiter:
 *        q = A p; den = p.q; if den <= 0 or non-finite: breakdown, iters = k;
 *        alpha = rho / den; x += alpha p; r -= alpha q;
 *        rel = ||r|| / bnorm; iters = k; if rel <= tol: converged;
 *        z = D^-1 r (same zero rule); rho' = r.z; beta = rho'/rho; p = z + beta p.
 *  - oracle_pcg_scipy: scipy.sparse.linalg.cg as called by solve_cpu_cg,
 *    code/fast_solver_backends.py:90-132, SciPy 1.15.3
 *    scipy/sparse/linalg/_isolve/iterative.py:375-423:
 *      bnrm2 = ||b||; if bnrm2 == 0 return x = b, 0 iterations;
 *      atol = rtol * bnrm2; r = b; for it = 0..maxiter-1:
 *        if ||r|| < atol: converged (strict); z = M r (zero diagonal -> 0);
 *        rho = r.z; p = (it > 0) ? z + (rho/rho_prev) p : z; q = A p;
 *        alpha = rho / p.q; x += alpha p; r -= alpha q; count++.
 *
 * Both return the iteration count and a status code using the same numbering
 * as the HIP library (1 converged, 2 breakdown, 3 maximum iterations).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_CONVERGED 1
#define OR_BREAKDOWN 2
#define OR_MAXITER 3

static void spmv(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals, const double* x,
                 double* y) {
  for (int32_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t j = indptr[i]; j < indptr[i + 1]; ++j) s += vals[j] * x[indices[j]];
    y[i] = s;
  }
}

static double dot(int32_t n, const double* a, const double* b) {
  double s = 0.0;
  for (int32_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static void diagonal(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals, double* d) {
  for (int32_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t j = indptr[i]; j < indptr[i + 1]; ++j)
      if (indices[j] == i) s += vals[j];
    d[i] = s;
  }
}

void oracle_spmv(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals, const double* x,
                 double* y) {
  spmv(n, indptr, indices, vals, x, y);
}

/* Reference GPU-loop semantics. iters_out/status_out/rel_out are scalars;
 * hist (may be NULL, size maxiter+1) receives ||r_k||/||b|| for k = 0..iters. */
int oracle_pcg_gpu_loop(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                        const double* b, double* x, double tol, int32_t maxiter, int32_t jacobi, int32_t* iters_out,
                        int32_t* status_out, double* rel_out, double* hist) {
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  double* d = (double*)malloc(sizeof(double) * (size_t)n);
  if (!r || !z || !p || !q || !d) {
    free(r); free(z); free(p); free(q); free(d);
    return -1;
  }
  diagonal(n, indptr, indices, vals, d);
  for (int32_t i = 0; i < n; ++i) {
    x[i] = 0.0;
    r[i] = b[i];
    z[i] = (jacobi && fabs(d[i]) > 0.0) ? r[i] / d[i] : r[i];
    p[i] = z[i];
  }
  double rho = dot(n, r, z);
  double bnorm = sqrt(dot(n, b, b));
  if (bnorm == 0.0) bnorm = 1.0;
  int32_t status = OR_MAXITER, iters = 0;
  double rel = sqrt(dot(n, r, r)) / bnorm;
  if (hist) hist[0] = rel;
  for (int32_t k = 1; k <= maxiter; ++k) {
    spmv(n, indptr, indices, vals, p, q);
    const double den = dot(n, p, q);
    if (den <= 0.0 || !isfinite(den)) {
      status = OR_BREAKDOWN;
      iters = k;
      break;
    }
    const double alpha = rho / den;
    for (int32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rel = sqrt(dot(n, r, r)) / bnorm;
    iters = k;
    if (hist) hist[k] = rel;
    if (rel <= tol) {
      status = OR_CONVERGED;
      break;
    }
    for (int32_t i = 0; i < n; ++i) z[i] = (jacobi && fabs(d[i]) > 0.0) ? r[i] / d[i] : r[i];
    const double rho_new = dot(n, r, z);
    const double beta = rho_new / rho;
    for (int32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rho = rho_new;
  }
  *iters_out = iters;
  *status_out = status;
  *rel_out = rel;
  free(r); free(z); free(p); free(q); free(d);
  return 0;
}

/* SciPy CG semantics (strict stop test at the loop top, Jacobi zero -> 0). */
int oracle_pcg_scipy(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals, const double* b,
                     double* x, double rtol, int32_t maxiter, int32_t jacobi, int32_t* iters_out,
                     int32_t* status_out) {
  const double bnrm2 = sqrt(dot(n, b, b));
  for (int32_t i = 0; i < n; ++i) x[i] = 0.0;
  if (bnrm2 == 0.0) {
    for (int32_t i = 0; i < n; ++i) x[i] = b[i];
    *iters_out = 0;
    *status_out = OR_CONVERGED;
    return 0;
  }
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  double* inv = (double*)malloc(sizeof(double) * (size_t)n);
  if (!r || !z || !p || !q || !inv) {
    free(r); free(z); free(p); free(q); free(inv);
    return -1;
  }
  diagonal(n, indptr, indices, vals, inv);
  for (int32_t i = 0; i < n; ++i) {
    inv[i] = fabs(inv[i]) > 0.0 ? 1.0 / inv[i] : 0.0;
    r[i] = b[i];
  }
  const double atol = rtol * bnrm2;
  double rho_prev = 0.0;
  int32_t count = 0, status = OR_MAXITER;
  for (int32_t it = 0; it < maxiter; ++it) {
    if (sqrt(dot(n, r, r)) < atol) {
      status = OR_CONVERGED;
      break;
    }
    for (int32_t i = 0; i < n; ++i) z[i] = jacobi ? inv[i] * r[i] : r[i];
    const double rho = dot(n, r, z);
    if (it > 0) {
      const double beta = rho / rho_prev;
      for (int32_t i = 0; i < n; ++i) p[i] = p[i] * beta + z[i];
    } else {
      memcpy(p, z, sizeof(double) * (size_t)n);
    }
    spmv(n, indptr, indices, vals, p, q);
    const double alpha = rho / dot(n, p, q);
    for (int32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rho_prev = rho;
    ++count;
  }
  *iters_out = count;
  *status_out = status;
  free(r); free(z); free(p); free(q); free(inv);
  return 0;
}
