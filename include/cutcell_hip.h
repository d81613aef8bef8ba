/*
 * cutcell_hip.h - C ABI of the MI355X (gfx950) cut-cell PCG library.
 *
 * All entry points are extern "C", take plain device pointers and sizes, run
 * asynchronously on the given HIP stream unless stated otherwise, and return
 * an int status: 0 on success, a negative CC_E* code on a bad argument, or
 * the positive hipError_t of a failed HIP call. Nothing throws across the ABI.
 *
 * Matrix format: CSR with int32 row pointers (n+1) and int32 column indices,
 * values in float64 (or float32 where the name says so). Multi-RHS blocks are
 * row-major [n][k] (the k right-hand sides of one row are contiguous), k in
 * {1, 2, 4, 8, 16}.
 *
 * Reference interfaces replaced (reference repository, file:line):
 *  - cc_pcg_jacobi_f64 / cc_pcg_jacobi_mixed: the Jacobi-PCG loop of
 *    solve_gpu_cg, code/fast_solver_backends.py:194-289 (its CuPy/cuSPARSE
 *    loop, lines 233-274), reached through solve_spd(backend="gpu"),
 *    code/fast_solver_backends.py:292-325;
 *  - cc_spmv_csr_* / cc_spmm_csr_*: the `A_gpu @ p` product of that loop,
 *    code/fast_solver_backends.py:250 (cuSPARSE csrmv through cupyx);
 *  - cc_residual_f64: relative_residual, code/fast_solver_backends.py:63-71;
 *  - cc_jacobi_diag_f64: the `A_gpu.diagonal()` of line 237 and the
 *    zero-diagonal-to-identity rule of line 238.
 */
#ifndef CUTCELL_HIP_H
#define CUTCELL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* cc_stream_t; /* a hipStream_t; NULL is the null stream */

enum {
  CC_OK = 0,
  CC_E_ARG = -1,       /* bad size / null pointer / unsupported k */
  CC_E_WORKSPACE = -2, /* workspace too small */
  CC_E_NODEVICE = -3,
  CC_E_TOO_LARGE = -4  /* the problem needs more than a fixed scratch holds */
};

/* Per-column solve status (the `message` of the reference report). */
enum {
  CC_RUNNING = 0,
  CC_CONVERGED = 1,   /* "converged" */
  CC_BREAKDOWN = 2,   /* "CG breakdown: non-positive curvature" */
  CC_MAXITER = 3      /* "maximum iterations reached" */
};

/* One right-hand side's outcome, written to host memory by the solvers. */
typedef struct cc_pcg_result {
  int32_t iterations;      /* CG iterations, counted as the reference GPU loop does */
  int32_t status;          /* CC_CONVERGED / CC_BREAKDOWN / CC_MAXITER */
  int32_t outer;           /* refinement sweeps (mixed precision), else 1 */
  int32_t inner_total;     /* total inner CG iterations (mixed precision) */
  double recursive_rel;    /* ||r_k|| / ||b|| of the recurrence at exit */
  double true_rel;         /* ||b - A x|| / ||b|| recomputed in FP64 on the device */
} cc_pcg_result;

/* Library / device information. */
int cc_version(void);
int cc_device_count(int* count);

/* Optional event timing of the solver kernels (no reference counterpart: it
 * serves the benchmark's live roofline figure). When enabled, every launch of
 * a timed kernel class inside the solvers is bracketed by two hipEvents on its
 * stream; cc_profile_read returns the summed milliseconds and launch count of
 * a class: 0 = SpMV/SpMM (or matrix-free apply) of the CG loops, 1 = CG
 * update, 2 = CG direction, 3 = batched persistent solve, 4 = learned
 * preconditioner apply. Process-wide and not thread-safe. Each timed launch
 * adds the dispatch of its two events to the stream, so cc_profile_enable_classes
 * times only the classes whose bit is set in mask (bit c = class c; 0 = off):
 * the benchmark brackets only the product it reports. cc_profile_enable(on)
 * times every class (on != 0) or none. */
int cc_profile_enable(int32_t on);
int cc_profile_enable_classes(int32_t mask);
int cc_profile_reset(void);
int cc_profile_read(int32_t which, double* total_ms, int64_t* count);
/* Folds the events of launches made outside the solvers' own polling (the
 * standalone SpMV/SpMM entry points and the batched solve) into the totals;
 * call after synchronising their stream. */
int cc_profile_collect(void);

/* SpMM/SpMV launch shape for k right-hand sides of value_bytes-wide vectors
 * (8 = float64, 4 = float32; each width keeps its own shape), chosen by the
 * autotuner (no reference counterpart): cols_per_lane right-hand sides per
 * lane, loaded as one vector (1, 2 or 4 with at most 16 bytes; 0 = widest),
 * nz_lanes nonzero slots per row team ((k/cols_per_lane)*nz_lanes <= 64,
 * power of two; 0 = from the mean row length), grid blocks (0 = one block per
 * few rows up to 1024), unroll 1, 4 or 8 (8 = software-pipelined: the next
 * entries' indices and values load while the current X rows are gathered;
 * 16 = for k = 1 only, the nonzero-window kernel: 4 entries per lane per step
 * with 16-byte loads and per-row sums of the products in LDS, used when
 * indices and vals are 16-byte aligned; 0 = 4), stream_loads 1 = nontemporal
 * loads of the matrix arrays (not with unroll 1), 0 = ordinary loads;
 * sweep_steps (0..64, 0 = 1): the rows are cut into grid * sweep_steps
 * chunks of equal nonzero counts and block b takes chunks b, b + grid, ...,
 * so that the grid streams one compact stretch of the matrix at a time. */
int cc_spmm_set_config(int32_t k, int32_t value_bytes, int32_t nz_lanes, int32_t grid, int32_t unroll,
                       int32_t cols_per_lane, int32_t stream_loads, int32_t sweep_steps);

/* y = A x, float64 (replaces the cuSPARSE product A_gpu @ p). */
int cc_spmv_csr_f64(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                    const double* x, double* y, cc_stream_t stream);
/* y = A x with float32 values and vectors (the inner product of the mixed solve). */
int cc_spmv_csr_f32(int32_t n, const int32_t* indptr, const int32_t* indices, const float* vals,
                    const float* x, float* y, cc_stream_t stream);
/* Y = A X for k right-hand sides, X and Y row-major [n][k]. */
int cc_spmm_csr_f64(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                    const double* X, double* Y, cc_stream_t stream);
int cc_spmm_csr_f32(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const float* vals,
                    const float* X, float* Y, cc_stream_t stream);

/* d[i] = A[i,i] (0 where the row stores no diagonal entry). */
int cc_jacobi_diag_f64(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                       double* diag, cc_stream_t stream);

/* Relative residuals ||B[:,c] - A X[:,c]|| / ||B[:,c]|| (||B[:,c]|| = 0 -> 1) for
 * the k columns, written to host memory rel[k]. Synchronises the stream. */
int cc_residual_f64(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                    const double* B, const double* X, double* rel, void* workspace, size_t workspace_bytes,
                    cc_stream_t stream);

/* Workspace needed by the solvers for n rows and k right-hand sides.
 * precision: 0 = float64 solve, 1 = mixed (float32 inner, float64 outer). */
size_t cc_pcg_workspace_bytes(int32_t n, int32_t k, int32_t precision);

/* Jacobi-preconditioned CG, float64 throughout, k independent right-hand
 * sides sharing A (replaces solve_gpu_cg's loop). X (row-major [n][k]) is
 * overwritten with the solutions; x0 = 0. Per column the recurrence, the
 * stopping test (||r_k||/||b|| <= tol, ||b|| = 0 -> 1), the breakdown test
 * (p.Ap <= 0 or non-finite) and the iteration count follow
 * code/fast_solver_backends.py:233-274. preconditioner: 0 = none, 1 = Jacobi
 * (zero diagonal -> identity). check_every: iterations between host polls of
 * the device status. history (device memory, may be NULL) receives
 * ||r_j||/||b|| of every iteration j = 0..iterations as history[j*k + c]
 * (size (maxiter+1)*k). Synchronises the stream before returning; results[k]
 * is host memory. */
int cc_pcg_jacobi_f64(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                      const double* B, double* X, double tol, int32_t maxiter, int32_t preconditioner,
                      int32_t check_every, void* workspace, size_t workspace_bytes, cc_pcg_result* results,
                      double* history, cc_stream_t stream);

/* Mixed precision: float32 values (vals32, same pattern) and vectors in the
 * inner Jacobi-PCG, float64 residual recomputation, dot products and outer
 * correction (iterative refinement). inner_tol is the inner solve's relative
 * tolerance; max_outer bounds the refinement sweeps. */
int cc_pcg_jacobi_mixed(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                        const float* vals32, const double* B, double* X, double tol, int32_t maxiter,
                        double inner_tol, int32_t max_outer, int32_t preconditioner, int32_t check_every,
                        void* workspace, size_t workspace_bytes, cc_pcg_result* results, cc_stream_t stream);

/* cc_pcg_jacobi_f64 without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. Same
 * arguments, kernels and results as cc_pcg_jacobi_f64; the workspace is a
 * fixed device array of cc_pcg_noalloc_max_bytes() (64 MiB) that the library
 * holds on every GPU it runs on until the process ends. Problems whose
 * cc_pcg_workspace_bytes(n, k, 0) exceeds it return CC_E_TOO_LARGE
 * (C2's n = 81,195 fits for every k up to 16). The entry point calls no
 * hipMalloc/hipFree. Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (several streams or host threads) race. The stream
 * must belong to the current device. */
size_t cc_pcg_noalloc_max_bytes(void);
int cc_pcg_jacobi_f64_noalloc(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices,
                              const double* vals, const double* B, double* X, double tol, int32_t maxiter,
                              int32_t preconditioner, int32_t check_every, cc_pcg_result* results, double* history,
                              cc_stream_t stream);

/* Restarted GMRES with the Jacobi preconditioner (CPU rule: zero diagonal
 * -> 0) on a CSR matrix, float64, one right-hand side, x0 = 0: the GMRES leg
 * of the solver interface (replaces the SciPy gmres call of solve_cpu_gmres,
 * code/fast_solver_backends.py:135-191, with its semantics: left
 * preconditioning, modified Gram-Schmidt, Givens rotations as LAPACK
 * dlartg, the inner stop on the preconditioned residual with the adaptive
 * tolerance, the outer stop on ||b - A x|| <= tol ||b||, maxiter counting
 * restarts (0 -> 10 n), restart (0 -> 20) capped at n). The vector work runs
 * on the device; the Hessenberg problem on the host, which reads one column
 * per inner step (so the call synchronises the stream each step). */
typedef struct cc_gmres_result {
  int32_t iterations;    /* inner steps, as SciPy's pr_norm callback counts them */
  int32_t info;          /* 0 converged, else maxiter */
  int32_t restarts;      /* outer cycles run */
  int32_t pad;
  double residual_norm;  /* ||b - A x|| at exit (float64, device) */
  double b_norm;         /* ||b|| */
} cc_gmres_result;
size_t cc_gmres_workspace_bytes(int32_t n, int32_t restart);
int cc_gmres_jacobi_f64(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                        const double* b, double* x, double tol, int32_t restart, int32_t maxiter,
                        int32_t preconditioner, void* workspace, size_t workspace_bytes, cc_gmres_result* result,
                        cc_stream_t stream);

/* Local Gram blocks of 3D level-set cut cells, the per-cut-cell integration
 * of the 3D matrix-free systems' assembly (done on the host by
 * cutcell/grid3d.py levelset_cut_data in a process pool; these geometries
 * have no reference code). Each cut cell's six Kuhn tetrahedra are clipped
 * to phi >= 0 of the linear interpolant of its corner values, split into
 * tetrahedra as the host splits them, and integrated by the collapsed Gauss
 * rule (rule_pts, rule_w: the host's tetrahedron rule); mass and stiff get the
 * (p+1)^3 x (p+1)^3 blocks, local function (iz, iy, ix) x fastest, and frac
 * the cell's fluid volume fraction. */
typedef struct cc_cut3d_desc {
  int32_t n_cut;
  int32_t degree;           /* 1 .. 3 */
  int32_t ne[3];            /* elements along x, y, z */
  int32_t n_knots[3];
  double length[3];
  const double* knots[3];   /* parametric knots per axis (x / length) */
  const int32_t* cells;     /* [n_cut][3]: ex, ey, ez */
  const double* corner;     /* [n_cut][8]: phi at corner c (x bit 0, y bit 1, z bit 2) */
  int32_t nq;
  const double* rule_pts;   /* [nq][3] */
  const double* rule_w;     /* [nq] */
} cc_cut3d_desc;
int cc_cut3d_gram(const cc_cut3d_desc* d, double* mass, double* stiff, double* frac, cc_stream_t stream);

/* 2D cut-cell assembly on the device (replaces the per-cell loops of
 * assemble_oblique_cut_matrices, code/oblique_cut_sbp.py:271-433, for the
 * mass H and stiffness K): N x N cells of degree-p tensor B-splines on open
 * uniform knots, the unit square cut by normal . x >= offset. */
typedef struct cc_cut2d_cell {
  int32_t ix, iy;            /* the cell */
  double normal_x, normal_y; /* the half plane of this cell's system */
  double offset;
} cc_cut2d_cell;
/* Local Gram blocks of ncells cut cells (any systems sharing N, p and q):
 * the square clipped against the half plane (Sutherland-Hodgman, tolerance
 * 1e-14, duplicates within 1e-13 dropped), fanned from its first vertex,
 * each triangle by the q x q collapsed Gauss rule of gauss_x / gauss_w
 * (Gauss-Legendre on [-1, 1]); H_blocks / K_blocks[cell][a][b] over the
 * (p+1)^2 local functions a = ay (p+1) + ax; areas[cell] = the clipped
 * area, or 0 when the polygon has fewer than 3 vertices or area <= 1e-16
 * (such a cell contributes nothing). degree 1..5. */
int cc_cut2d_gram(int32_t ncells, const cc_cut2d_cell* cells, int32_t n_elements, int32_t degree, int32_t q,
                  const double* gauss_x, const double* gauss_w, double* H_blocks, double* K_blocks, double* areas,
                  cc_stream_t stream);
/* Row counts and row pointers (n = (N+p)^2 + 1) of H and K for one system:
 * entry (i, j) exists when a contributing cell (cell_class[iy * N + ix]: 0
 * empty, 1 full, 2 cut, 3 cut but contributing nothing) supports both i and
 * j. */
int cc_asm2d_count(int32_t n_elements, int32_t degree, const int8_t* cell_class, int32_t* row_counts,
                   int32_t* indptr, cc_stream_t stream);
/* Columns (ascending) and values of H and K: full cells from the 1D element
 * forms elem_mass / elem_stiff ([N][p+1][p+1], the same basis in x and y),
 * cut cells from the blocks at cut_slot[cell]; each entry summed over its
 * covering cells in a fixed order (no atomics). */
int cc_asm2d_fill(int32_t n_elements, int32_t degree, const int8_t* cell_class, const int32_t* cut_slot,
                  const double* elem_mass, const double* elem_stiff, const double* H_blocks, const double* K_blocks,
                  const int32_t* indptr, int32_t* indices, double* H_vals, double* K_vals, cc_stream_t stream);

/* Row-block form of a CSR matrix (no reference counterpart: a derived layout
 * for products with several right-hand sides, built once per matrix). Rows
 * are taken in groups of CC_RB_ROWS consecutive rows; group g stores the
 * ascending union of its rows' columns, gcol[gptr[g] .. gptr[g+1]), and per
 * union column u the CC_RB_ROWS row values gval[u * CC_RB_ROWS + r] (zero
 * where row r has no entry at that column), so one gathered row of X serves
 * all rows of the group. ngroups = ceil(n / CC_RB_ROWS). */
#define CC_RB_ROWS 4
typedef struct cc_rb_desc {
  int32_t n;
  int32_t ngroups;
  const int32_t* gptr;  /* [ngroups + 1] */
  const int32_t* gcol;  /* [nu] */
  const double* gval;   /* [nu * CC_RB_ROWS] float64 values, or NULL */
  const float* gval32;  /* [nu * CC_RB_ROWS] float32 values, or NULL */
  /* Optional (both NULL, or both set): the union columns of each group as
   * runs of consecutive columns (cc_runs_count / cc_runs_fill with units =
   * groups, uptr = gptr, cols = gcol): grptr [ngroups + 1], grun
   * [2 * nruns]. When set, products whose group team is one wave read the
   * columns from the runs instead of gcol (same values, same sums). */
  const int32_t* grptr;
  const uint32_t* grun;
  /* Optional: paired groups. When set, group g holds the 8 rows of the two
   * 4-row groups pairs[2g] and pairs[2g+1] (the second -1 for a group left
   * alone; every 4-row group appears exactly once), ngroups is the number of
   * pairs, and gval / gval32 hold 8 values per union column (built by
   * cc_rb_build_pairs_count / _fill; grptr, grun must then be NULL). */
  const int32_t* pairs;
} cc_rb_desc;

/* Build step 1: from a CSR matrix whose column indices ascend within each
 * row (repeated columns are summed), write the group offsets to gptr
 * (device, ngroups + 1 entries) and the union size to *nu (host).
 * Synchronises the stream; returns CC_E_ARG if a row's indices descend. */
int cc_rb_build_count(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* gptr, int64_t* nu,
                      cc_stream_t stream);
/* Build step 2: fill gcol [nu] and gval / gval32 [nu * CC_RB_ROWS] (either
 * may be NULL, not both) from the same CSR matrix and the gptr of step 1. */
int cc_rb_build_fill(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                     const int32_t* gptr, int32_t* gcol, double* gval, float* gval32, cc_stream_t stream);
/* The same two build steps for paired groups (pairs [npairs * 2], see
 * cc_rb_desc.pairs): the union of each pair's 8 rows, 8 values per union
 * column. */
int cc_rb_build_pairs_count(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t npairs,
                            const int32_t* pairs, int32_t* gptr, int64_t* nu, cc_stream_t stream);
int cc_rb_build_pairs_fill(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                           int32_t npairs, const int32_t* pairs, const int32_t* gptr, int32_t* gcol, double* gval,
                           float* gval32, cc_stream_t stream);
/* Launch shape of the row-block product for k right-hand sides of
 * value_bytes-wide vectors (autotuner hook): nz_lanes union-column slots per
 * group team ((k/cols_per_lane)*nz_lanes <= 64, power of two, 0 = default),
 * grid blocks (0 = default, at most 1024), unroll 1 or 4 (0 = 4),
 * cols_per_lane 1, 2 or 4 right-hand sides per lane (at most 16 bytes,
 * 0 = widest). */
int cc_rb_set_config(int32_t k, int32_t value_bytes, int32_t nz_lanes, int32_t grid, int32_t unroll,
                     int32_t cols_per_lane);
/* Y = A X for k in {1,2,4,8,16} right-hand sides (row-major [n][k]) from the
 * row-block form (the same product as cc_spmm_csr_f64 / _f32). */
int cc_spmm_rb_f64(const cc_rb_desc* rb, int32_t k, const double* X, double* Y, cc_stream_t stream);
int cc_spmm_rb_f32(const cc_rb_desc* rb, int32_t k, const float* X, float* Y, cc_stream_t stream);
/* Run encoding of sorted column lists (no reference counterpart; it
 * replaces the index stream of the cuSPARSE product A_gpu @ p,
 * code/fast_solver_backends.py:250). A unit u (a CSR row) holds entries
 * uptr[u] .. uptr[u+1]-1 of cols, strictly ascending; it is stored as its
 * runs of consecutive columns, two uint32 per run: the first column, and
 * (offset of the run's first entry within the unit) << 8 | length (1..255).
 * cc_runs_count writes the run offsets rptr [nunits + 1] (device) and the
 * total to *nruns (host; synchronises; CC_E_ARG if a unit's columns do not
 * ascend strictly or it has 2^24 entries or more); cc_runs_fill writes the
 * runs [2 * nruns]. */
int cc_runs_count(int32_t nunits, const int32_t* uptr, const int32_t* cols, int32_t* rptr, int64_t* nruns,
                  cc_stream_t stream);
int cc_runs_fill(int32_t nunits, const int32_t* uptr, const int32_t* cols, const int32_t* rptr, uint32_t* runs,
                 cc_stream_t stream);
/* y = A x from the runs of A's rows and its CSR values (indptr, vals): the
 * same product as cc_spmv_csr_f64 / _f32 (code/fast_solver_backends.py:250),
 * each row's sum taken run by run. */
int cc_spmv_runs_f64(int32_t n, const int32_t* rptr, const uint32_t* runs, const int32_t* indptr, const double* vals,
                     const double* x, double* y, cc_stream_t stream);
int cc_spmv_runs_f32(int32_t n, const int32_t* rptr, const uint32_t* runs, const int32_t* indptr, const float* vals,
                     const float* x, float* y, cc_stream_t stream);
/* cc_pcg_jacobi_f64 / cc_pcg_jacobi_mixed for one right-hand side with the
 * CG products y = A p taken from the runs of A's rows (rptr, runs of
 * cc_runs_fill) and its CSR values: the reference's GPU loop
 * (code/fast_solver_backends.py:233-274) with its product
 * (code/fast_solver_backends.py:250) from the run encoding. The diagonal and
 * the float64 true residual come from the CSR arrays. Workspace as
 * cc_pcg_workspace_bytes(n, 1, precision). */
int cc_pcg_jacobi_f64_runs(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                           const int32_t* rptr, const uint32_t* runs, const double* B, double* X, double tol,
                           int32_t maxiter, int32_t preconditioner, int32_t check_every, void* workspace,
                           size_t workspace_bytes, cc_pcg_result* results, double* history, cc_stream_t stream);
int cc_pcg_jacobi_mixed_runs(int32_t n, const int32_t* indptr, const int32_t* indices, const double* vals,
                             const float* vals32, const int32_t* rptr, const uint32_t* runs, const double* B,
                             double* X, double tol, int32_t maxiter, double inner_tol, int32_t max_outer,
                             int32_t preconditioner, int32_t check_every, void* workspace, size_t workspace_bytes,
                             cc_pcg_result* results, cc_stream_t stream);
/* cc_pcg_jacobi_f64 / cc_pcg_jacobi_mixed with the CG products taken from
 * the row-block form rb of the same matrix (rb->gval for float64,
 * rb->gval32 for the mixed inner solve); the diagonal and the float64 true
 * residuals still come from the CSR arrays. Same workspace sizes. */
int cc_pcg_jacobi_f64_rb(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                         const cc_rb_desc* rb, const double* B, double* X, double tol, int32_t maxiter,
                         int32_t preconditioner, int32_t check_every, void* workspace, size_t workspace_bytes,
                         cc_pcg_result* results, double* history, cc_stream_t stream);
int cc_pcg_jacobi_mixed_rb(int32_t n, int32_t k, const int32_t* indptr, const int32_t* indices, const double* vals,
                           const cc_rb_desc* rb, const double* B, double* X, double tol, int32_t maxiter,
                           double inner_tol, int32_t max_outer, int32_t preconditioner, int32_t check_every,
                           void* workspace, size_t workspace_bytes, cc_pcg_result* results, cc_stream_t stream);

/* The aggregated operator E^T B E applied without forming it (replaces the
 * triple product restrict_spd, code/bell_3d_core.py:550-552, and the cuSPARSE
 * product A_gpu @ p by it, code/fast_solver_backends.py:250). B is the
 * background operator (K + H, symmetrised) over the active dofs, its rows and
 * columns renumbered: index j < n_red is the interior dof of reduced unknown
 * j, index n_red + s is small dof s. E_s (n_small rows) holds the small dofs'
 * extension weights over the reduced unknowns; T (n_touch rows) is E_s^T
 * restricted to the reduced unknowns it reaches (t_rows ascending, each row's
 * small dofs t_small with weights t_vals). One apply is three launches on a
 * combined vector [x ; u_s] of n_red + n_small rows: the small dofs' gather
 * u_s = E_s x, the product [y ; v_s] = B [x ; u_s] (the autotuned CSR
 * product, or the row-block one when b_rb is set), and y[t_rows] += T v_s
 * summed in a fixed order. diag (may be NULL) is the Jacobi diagonal of
 * E^T B E. All arrays are device memory. */
typedef struct cc_ekte_desc {
  int32_t n_red;
  int32_t n_small;
  int32_t n_touch;
  int32_t pad;
  int64_t nnz_b;            /* nonzeros of B (picks the default product shape) */
  const int32_t* b_indptr;  /* [n_red + n_small + 1] */
  const int32_t* b_indices; /* combined numbering */
  const double* b_vals;
  const int32_t* e_indptr;  /* [n_small + 1] */
  const int32_t* e_indices; /* reduced unknowns */
  const double* e_vals;
  const int32_t* t_rows;    /* [n_touch] reduced unknowns */
  const int32_t* t_indptr;  /* [n_touch + 1] */
  const int32_t* t_small;   /* small dof indices */
  const double* t_vals;
  const struct cc_rb_desc* b_rb; /* row-block form of B (float64 values), or NULL for the CSR product */
  const double* diag;       /* [n_red] or NULL */
} cc_ekte_desc;

/* Workspace of one apply (solve = 0) or of cc_pcg_ekte (solve = 1) on k
 * right-hand sides; 0 on a bad descriptor or k. */
size_t cc_ekte_workspace_bytes(const cc_ekte_desc* d, int32_t k, int32_t solve);
/* Y = E^T B E X for k in {1, 2, 4, 8, 16} columns (row-major [n_red][k]);
 * X is copied into and Y out of the workspace's combined vectors. */
int cc_ekte_apply(const cc_ekte_desc* d, int32_t k, const double* X, double* Y, void* workspace,
                  size_t workspace_bytes, cc_stream_t stream);
/* Jacobi-PCG (float64, the reference GPU loop's recurrence and stopping
 * rules, code/fast_solver_backends.py:233-274) on the fused operator for k
 * right-hand sides; results[k] host memory; synchronises the stream. Its
 * direction and product vectors carry the n_small tail rows, so the products
 * inside the loop copy nothing. */
int cc_pcg_ekte(const cc_ekte_desc* d, int32_t k, const double* B, double* X, double tol, int32_t maxiter,
                int32_t preconditioner, int32_t check_every, void* workspace, size_t workspace_bytes,
                cc_pcg_result* results, cc_stream_t stream);

/* Batched persistent solve of nsys independent systems of uneven size in one
 * launch (replaces a loop of solve_spd calls over independent cases, the
 * reference's case-level parallelism in code/fast_solver_workbench.py:201-217,
 * each case being solve_gpu_cg's loop, code/fast_solver_backends.py:233-274).
 * The systems' CSR arrays are concatenated: system s has sys_n[s] rows, its
 * row pointers (local, starting at 0) at indptr + sys_indptr_off[s], its
 * column indices (local) and values at sys_nnz_off[s], and its right-hand side
 * and solution at sys_vec_off[s]. All per-system arrays and order (the
 * dispatch order, longest first) are device memory. n_max <= 8192 (the
 * workgroup keeps p and Ap of its system in LDS) and must be at least
 * max(sys_n): the launch shape is chosen from n_max, and a system with more
 * rows than that shape holds gets status CC_E_ARG (-1) in its result, with
 * its solution left untouched. precision 0 = float64,
 * 1 = float32 inner / float64 outer (vals32 required). block_threads in
 * {256, 512, 1024} and team_lanes in {4, 8, 16, 32, 64} come from the
 * autotuner; grid <= 0 means one workgroup per system. counter is one device
 * int32 of scratch; results[nsys] is device memory. Asynchronous. */
int cc_pcg_batched(int32_t nsys, int32_t n_max, const int32_t* sys_n, const int64_t* sys_indptr_off,
                   const int64_t* sys_nnz_off, const int64_t* sys_vec_off, const int32_t* order,
                   const int32_t* indptr, const int32_t* indices, const double* vals, const float* vals32,
                   const double* B, double* X, double tol, int32_t maxiter, int32_t preconditioner,
                   int32_t precision, double inner_tol, int32_t max_outer, int32_t block_threads,
                   int32_t team_lanes, int32_t grid, int32_t* counter, cc_pcg_result* results,
                   cc_stream_t stream);

/* Matrix-free aggregated operator y = E^T (K + c M) E x on a tensor B-spline
 * background grid (no reference counterpart in code: the reference only ever
 * forms E^T A E explicitly, restrict_spd in code/bell_3d_core.py:550-552, and
 * multiplies it with cuSPARSE, code/fast_solver_backends.py:250; this is the
 * same operator applied without the matrix). All pointers are device memory;
 * the struct itself is host memory. Dof order is x fastest; on element e the
 * live functions are e + a (a = 0..degree), modulo nb on a periodic axis.
 *  - band_m/band_k[axis]: global 1D mass/stiffness, [nb][2*degree+1], entry
 *    (i, i + d - degree) at [i][d];
 *  - elem_m/elem_k[axis]: per-element 1D forms, [ne][degree+1][degree+1];
 *  - corr_cell[n_corr][3] (ex, ey, ez): the shell of cells that are not
 *    fully fluid but touch an active dof; corr_block[s] >= 0 indexes
 *    blocks[][P][P] (P = (degree+1)^3, local order (z, y, x)), the cut cell's
 *    Gram block of K + c M minus the full cell's tensor block (C_e - A_e,
 *    the correction the cell adds to the full-grid tensor apply), and -1
 *    marks an empty cell (correction -A_e, applied by sum factorisation);
 *  - touch_dof/touch_ptr/touch_idx: CSR over the dofs that receive shell
 *    corrections, entries index the per-cell outputs (s * P + local);
 *  - red2full[n_red]: interior dof of each unknown; full2red/full2small:
 *    inverse maps (-1 elsewhere); small_ptr/small_col/small_w: rows of E for
 *    the n_small small dofs (small_dof), columns are unknowns; et_ptr/
 *    et_small/et_w: the same entries grouped by unknown (E^T), et_rows the
 *    unknowns that have any;
 *  - diag[n_red]: diag(E^T A E) for the Jacobi preconditioner. */
typedef struct cc_sumfact_desc {
  int32_t nb[3];
  int32_t ne[3];
  int32_t degree;
  int32_t periodic[3];
  double mass_coef;
  const double* band_m[3];
  const double* band_k[3];
  const double* elem_m[3];
  const double* elem_k[3];
  int32_t n_corr;
  const int32_t* corr_cell;
  const int32_t* corr_block;
  const double* blocks;
  int32_t n_touch;
  const int32_t* touch_dof;
  const int32_t* touch_ptr;
  const int32_t* touch_idx;
  int32_t n_red;
  int32_t n_small;
  const int32_t* red2full;
  const int32_t* full2red;
  const int32_t* full2small;
  const int32_t* small_ptr;
  const int32_t* small_col;
  const double* small_w;
  const int32_t* small_dof;
  const int32_t* et_ptr;
  const int32_t* et_small;
  const double* et_w;
  const double* diag;
  int32_t n_et_rows;      /* unknowns with E^T entries (et_ptr[j + 1] > et_ptr[j]) */
  const int32_t* et_rows; /* their indices, ascending */
} cc_sumfact_desc;

/* Workspace of cc_sumfact_apply (and, with solve != 0, of cc_pcg_sumfact). */
size_t cc_sumfact_workspace_bytes(const cc_sumfact_desc* d, int32_t solve);
/* y = E^T (K + c M) E x, float64; degree 2 or 3. Asynchronous. */
int cc_sumfact_apply(const cc_sumfact_desc* d, const double* x, double* y, void* workspace, size_t workspace_bytes,
                     cc_stream_t stream);
/* Jacobi-PCG (reference GPU-loop semantics, as cc_pcg_jacobi_f64 with k = 1)
 * on the matrix-free operator. Synchronises; result is host memory; history
 * as in cc_pcg_jacobi_f64 (may be NULL). snapshots (device, [n_snap][n_red],
 * may be NULL when n_snap = 0) receives the iterate x_j after snap_iters[j]
 * (host array) iterations, or the final iterate where the solve stopped
 * earlier: the (residual, error) pairs of the learned preconditioner's
 * training set come from these iterates. */
int cc_pcg_sumfact(const cc_sumfact_desc* d, const double* b, double* x, double tol, int32_t maxiter,
                   int32_t preconditioner, int32_t check_every, void* workspace, size_t workspace_bytes,
                   cc_pcg_result* result, double* history, const int32_t* snap_iters, int32_t n_snap,
                   double* snapshots, cc_stream_t stream);

/* The same operator for k in {1, 2, 4, 8, 16} right-hand sides at once (X,
 * Y, B row-major [n_red][k]); each cut block is streamed from memory once per
 * apply for all k columns. cc_pcg_sumfact_k runs Jacobi-PCG on the k columns (semantics of
 * cc_pcg_jacobi_f64 per column; no history or snapshots); synchronises;
 * results[k] host memory. (No reference counterpart: the reference solves
 * one right-hand side per call, code/fast_solver_backends.py:194-289.) */
size_t cc_sumfact_k_workspace_bytes(const cc_sumfact_desc* d, int32_t k, int32_t solve);
int cc_sumfact_apply_k(const cc_sumfact_desc* d, int32_t k, const double* X, double* Y, void* workspace,
                       size_t workspace_bytes, cc_stream_t stream);
int cc_pcg_sumfact_k(const cc_sumfact_desc* d, int32_t k, const double* B, double* X, double tol, int32_t maxiter,
                     int32_t preconditioner, int32_t check_every, void* workspace, size_t workspace_bytes,
                     cc_pcg_result* results, cc_stream_t stream);

/* FP32 inner / FP64 outer Jacobi-PCG on the matrix-free operator for k in
 * {1, 2, 4, 8, 16} right-hand sides: the inner solves apply the operator in
 * float32 (cut-cell corrections from blocks32, a float32 copy of d->blocks
 * in the same layout; vectors, the tensor product and the shell in float32)
 * to inner_tol, and the outer loop recomputes B - A X with the float64
 * operator and stops on that residual (<= tol), at most max_outer sweeps.
 * results[c].iterations counts inner iterations, .outer the sweeps.
 * Replaces the loop of solve_gpu_cg (code/fast_solver_backends.py:233-274) on
 * E^T A E (code/bell_3d_core.py:550-552) with a mixed-precision apply; the
 * reference has no counterpart for the precision split.
 * cc_sumfact_apply_k_f32: Y = A X in float32 (workspace as
 * cc_sumfact_k_workspace_bytes(d, k, 0)). */
size_t cc_sumfact_mixed_k_workspace_bytes(const cc_sumfact_desc* d, int32_t k);
int cc_sumfact_apply_k_f32(const cc_sumfact_desc* d, const float* blocks32, int32_t k, const float* X, float* Y,
                           void* workspace, size_t workspace_bytes, cc_stream_t stream);
int cc_pcg_sumfact_mixed_k(const cc_sumfact_desc* d, const float* blocks32, int32_t k, const double* B, double* X,
                           double tol, int32_t maxiter, double inner_tol, int32_t max_outer, int32_t preconditioner,
                           int32_t check_every, void* workspace, size_t workspace_bytes, cc_pcg_result* results,
                           cc_stream_t stream);

/* ---- Learned preconditioner (no reference counterpart: the reference has
 * only Jacobi, code/fast_solver_backends.py:74-87 and :237-238). ---------
 * Activations are channel-last bfloat16 grids [voxel][16] on the background
 * dof grid (nx, ny, nz; x fastest); weights are bfloat16 [27][16][16]
 * (tap t = (dz+1)*9 + (dy+1)*3 + (dx+1), then input channel, then output
 * channel); mask[voxel] != 0 marks the voxels that carry an unknown. */
typedef struct cc_cnn_desc {
  int32_t nx, ny, nz;
  int32_t periodic[3];
  const uint16_t* weights[4]; /* the four layers: W0, W1, flip W1, flip W0 */
  const uint8_t* mask;
  const int32_t* red2full; /* voxel of each unknown */
  int32_t n_red;
} cc_cnn_desc;

/* D = A B for 16x16 bfloat16 A, B (row-major), fp32 D: the MFMA lane-layout
 * probe the convolution kernels rely on. */
int cc_mfma_probe(const uint16_t* A, const uint16_t* B, float* D, cc_stream_t stream);
/* Launch shape of the convolutions (autotuner hook, no reference
 * counterpart): work items (one grid of a batch times one z-tile of
 * 16 x 4 x 4 voxels) run in turn by one block, 1..4096; 0 = about 1024
 * blocks in all. */
int cc_cnn_set_config(int32_t items_per_block);

/* out = mask * act(conv3(in, weights)) [* (gate > 0)]; out (bfloat16) and/or
 * out_f32 (fp32) may be given. relu: 1 = ReLU. gate may be NULL. */
int cc_cnn_conv3_forward(int32_t nx, int32_t ny, int32_t nz, const int32_t* periodic, const uint16_t* in,
                         const uint16_t* weights, const uint8_t* mask, const uint16_t* gate, int32_t relu,
                         uint16_t* out, float* out_f32, cc_stream_t stream);
/* Weights of the data-gradient layer: taps reversed, channels transposed. */
int cc_cnn_flip_weights(const uint16_t* weights, uint16_t* flipped, cc_stream_t stream);
/* dW (fp32 [27][16][16]) = scale * sum_v in[v + off_t][ci] gout[v][co]
 * (+ dW when accumulate != 0), reduced over `blocks` partials in a fixed
 * order. */
size_t cc_cnn_wgrad_workspace_bytes(int32_t nx, int32_t ny, int32_t nz, int32_t blocks);
int cc_cnn_conv3_wgrad(int32_t nx, int32_t ny, int32_t nz, const int32_t* periodic, const uint16_t* in,
                       const uint16_t* gout, float* dW, float scale, int32_t accumulate, int32_t blocks,
                       void* workspace, size_t workspace_bytes, cc_stream_t stream);
/* The preconditioner network is symmetric: C = L1 L0 (masked linear layers,
 * W0 lifting channel 0 to 16 channels, W1 mixing 16 -> 16) and
 *     z = D^-1 r + S C^T C S r,  S = D^-1/2,
 * evaluated as four layers with weights (W0, W1, flip W1, flip W0).
 * cc_cnn_sym_weights writes those four bf16 layers from the fp32 masters
 * [2][27][16][16]; cc_cnn_sym_grads turns the four layers' gradients
 * [4][27][16][16] into the masters' (g0 + flip g3, g1 + flip g2), added to
 * g_master when accumulate != 0. */
int cc_cnn_sym_weights(const float* w_master, uint16_t* layers, cc_stream_t stream);
int cc_cnn_sym_grads(const float* g_layers, float* g_master, int32_t accumulate, cc_stream_t stream);
/* One Adam step on fp32 master weights w (n values); also refreshes the
 * bfloat16 copy. frozen (may be NULL) marks entries kept fixed. */
int cc_cnn_adam(int32_t n, float* w, const float* grad, float* m1, float* m2, float lr, float beta1, float beta2,
                float eps, int32_t step, const uint8_t* frozen, uint16_t* w_bf16, cc_stream_t stream);
/* z = D^-1 r + S y, y = channel 0 of the network applied to S r (S = D^-1/2;
 * a non-positive diagonal entry counts as 1), float64 in/out. */
size_t cc_cnn_precondition_workspace_bytes(const cc_cnn_desc* net);
int cc_cnn_precondition(const cc_cnn_desc* net, const double* r, const double* diag, double* z, void* workspace,
                        size_t workspace_bytes, cc_stream_t stream);
/* Training helpers: the network input scale * S r as channel 0 of a bf16
 * grid (other entries untouched), and the loss ||z - scale e||^2 /
 * ||scale e||^2 with z = scale D^-1 r + S y (y: fp32 channel-0 output), its
 * gradient dL/dy (times gscale) written to channel 0 of gout, and per-block
 * partials (||z - t||^2, ||t||^2) in partial[2 * nparts]. r, e: float32. */
int cc_cnn_train_input(int32_t n, const int32_t* red2full, const float* r, const double* diag, float scale,
                       uint16_t* grid, cc_stream_t stream);
int cc_cnn_train_loss(int32_t n, const int32_t* red2full, const float* r, const float* e, const double* diag,
                      float scale, const float* y, float inv_den, float gscale, uint16_t* gout, double* partial,
                      int32_t* nparts, cc_stream_t stream);
/* Flexible CG (one stored direction) with the learned preconditioner on the
 * matrix-free operator; stopping test, iteration count, breakdown test and
 * float64 true residual as in cc_pcg_sumfact. Synchronises. */
size_t cc_fcg_sumfact_workspace_bytes(const cc_sumfact_desc* d, const cc_cnn_desc* net);
int cc_fcg_sumfact(const cc_sumfact_desc* d, const cc_cnn_desc* net, const double* b, double* x, double tol,
                   int32_t maxiter, int32_t check_every, void* workspace, size_t workspace_bytes,
                   cc_pcg_result* result, double* history, cc_stream_t stream);

/* ---- Multilevel learned preconditioner (no reference counterpart). ------
 * Level 0 is the background dof grid; level l + 1 keeps every other grid
 * point of level l (clamped axis: n_c = (n_f + 1) / 2, coarse c on fine 2c;
 * periodic axis: n_f even, n_c = n_f / 2). P is tensor-product linear
 * interpolation and the restriction is exactly R = P^T. With Q_l = P_1 ... P_l
 * and C_l the symmetric network of level l (four layers as above),
 *     z = D^-1 r + S sum_{l with a net} Q_l C_l^T C_l Q_l^T S r,
 * symmetric and never below Jacobi. net[l] gives every level's grid and
 * periodic flags (periodic flags equal on all levels); levels with has_net[l]
 * also give weights and mask; net[0].red2full / n_red are the unknowns. The
 * coarsest level must carry a network. */
#define CC_ML_MAX_LEVELS 8
typedef struct cc_mlcnn_desc {
  int32_t levels;
  int32_t has_net[CC_ML_MAX_LEVELS];
  cc_cnn_desc net[CC_ML_MAX_LEVELS];
} cc_mlcnn_desc;

/* coarse = R fine, fine = P coarse (+ add) on single-channel float32 grids of
 * dims nf[3] -> nc[3] (x fastest), elements read/written with the given
 * strides (16 reads channel 0 of a [voxel][16] tensor). */
int cc_mg_restrict(const int32_t* nf, const int32_t* nc, const int32_t* periodic, const float* fine,
                   int32_t fine_stride, float* coarse, int32_t coarse_stride, cc_stream_t stream);
int cc_mg_prolong(const int32_t* nf, const int32_t* nc, const int32_t* periodic, const float* coarse,
                  int32_t coarse_stride, const float* add, int32_t add_stride, float* fine, cc_stream_t stream);
/* grid[v][0] = bf16(src[v * src_stride]) for every voxel (other channels untouched). */
int cc_mg_to_ch0(int64_t nvox, const float* src, int32_t src_stride, uint16_t* grid, cc_stream_t stream);
/* Training helpers of the multilevel network: u0[voxel of j] = scale * S r_j
 * (float32, other voxels untouched), and the loss of cc_cnn_train_loss with
 * the summed correction t0 (single-channel float32 on the background grid)
 * in place of y; its gradient dL/dt0 (times gscale) goes to g0 (float32). */
int cc_mg_train_input(int32_t n, const int32_t* red2full, const float* r, const double* diag, float scale, float* u0,
                      cc_stream_t stream);
int cc_mg_train_loss(int32_t n, const int32_t* red2full, const float* r, const float* e, const double* diag,
                     float scale, const float* t0, float inv_den, float gscale, float* g0, double* partial,
                     int32_t* nparts, cc_stream_t stream);
/* z = M r with the multilevel preconditioner, float64 in/out. */
size_t cc_mlcnn_precondition_workspace_bytes(const cc_mlcnn_desc* ml);
int cc_mlcnn_precondition(const cc_mlcnn_desc* ml, const double* r, const double* diag, double* z, void* workspace,
                          size_t workspace_bytes, cc_stream_t stream);
/* Flexible CG with the multilevel preconditioner on the matrix-free
 * operator; semantics as cc_fcg_sumfact. Synchronises. */
size_t cc_fcg_sumfact_ml_workspace_bytes(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml);
int cc_fcg_sumfact_ml(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml, const double* b, double* x, double tol,
                      int32_t maxiter, int32_t check_every, void* workspace, size_t workspace_bytes,
                      cc_pcg_result* result, double* history, cc_stream_t stream);
/* The same on k in {1,2,4,8,16} right-hand sides at once (B, X row-major
 * [n_red][k], results[k]): one K-column operator apply and one batched
 * network pass per iteration serve all columns; each column follows its own
 * flexible-CG recurrence and stops on its own (a zero column converges at
 * iteration 0). No reference counterpart (the reference solves one
 * right-hand side per call, code/fast_solver_backends.py:194-289).
 * Synchronises. */
size_t cc_fcg_sumfact_ml_k_workspace_bytes(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml, int32_t k);
int cc_fcg_sumfact_ml_k(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml, int32_t k, const double* B, double* X,
                        double tol, int32_t maxiter, int32_t check_every, void* workspace, size_t workspace_bytes,
                        cc_pcg_result* results, cc_stream_t stream);

/* The batched learned flexible CG of cc_fcg_sumfact_ml_k inside float64
 * iterative refinement: each sweep solves A z = r / ||r|| to inner_tol by
 * flexible CG whose operator applies are float32 (blocks32 as in
 * cc_pcg_sumfact_mixed_k; directions rounded to float32, products widened
 * back), then x += ||r|| z and r = B - A x in float64; stops on that residual
 * (<= tol), at most max_outer sweeps. results[c].iterations counts inner
 * iterations, .outer the sweeps. No reference counterpart (the reference has
 * Jacobi only, code/fast_solver_backends.py:74-87). */
size_t cc_fcg_sumfact_ml_mixed_k_workspace_bytes(const cc_sumfact_desc* d, const cc_mlcnn_desc* ml, int32_t k);
int cc_fcg_sumfact_ml_mixed_k(const cc_sumfact_desc* d, const float* blocks32, const cc_mlcnn_desc* ml, int32_t k,
                              const double* B, double* X, double tol, int32_t maxiter, double inner_tol,
                              int32_t max_outer, int32_t check_every, void* workspace, size_t workspace_bytes,
                              cc_pcg_result* results, cc_stream_t stream);

/* ---- Network pass of the multilevel preconditioner on gfx950 (no
 * reference counterpart). The preconditioners above evaluate every level's
 * C^T C in three launches that serve all levels and all grids of a batch:
 * the first two layers fused (the W0 layer computed on each tile's halo in
 * LDS), the flip W1 layer, and the flip W0 layer written as channel 0 in
 * float32; all on the 16x16x32 bf16 MFMA with fp32 accumulation.
 * cc_mfma32_probe: D = A B for bf16 A (16 x 32) and B (32 x 16), row-major,
 * fp32 D (16 x 16), the lane layout those kernels rely on.
 * cc_net_set_config: most blocks per pass launch (0 = two per CU of the
 * current device, all resident at once) and
 * legacy = 1 to run the four per-level layer launches of
 * cc_cnn_conv3_forward instead.
 * cc_net_level_forward: y = C^T C u (channel 0, masked) on nb grids of one
 * level, u and y float32 [nb][voxel]. */
int cc_mfma32_probe(const uint16_t* A, const uint16_t* B, float* D, cc_stream_t stream);
int cc_net_set_config(int32_t blocks, int32_t legacy);
size_t cc_net_level_workspace_bytes(const cc_cnn_desc* net, int32_t nb);
int cc_net_level_forward(const cc_cnn_desc* net, int32_t nb, const float* u, float* y, void* workspace,
                         size_t workspace_bytes, cc_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CUTCELL_HIP_H */
