/*
 * laamd: MI355X (gfx950) sparse hot path behind a C ABI.
 *
 * CSR SpMV (row-parallel, segmented-sum and ALBUS nnz-balanced forms),
 * level-scheduled and sync-free sparse triangular solves, the ILU(k)/IC(k)
 * preconditioner apply built on them, and the PCG / FGMRES loops that drive
 * them, with FP64 and mixed FP32/BF16 precision. Host-side symbolic work
 * (ILU(k)/IC(k) fill, level sets, RCM) is plain C++ in the same library and
 * needs no GPU.
 *
 * Conventions (mirroring the reference's free functions, which take raw
 * pointers and an explicit base, sparse_mat_op/spmv.hpp:16-19):
 *  - indices are int32, `base` is 0 or 1 and applies to ai and aj;
 *  - host-side functions take host pointers; functions on la_mat_t /
 *    la_trsv_t / la_prec_t handles take DEVICE pointers for vectors;
 *  - every function returns an la_status; 0 = LA_OK. Nothing throws across
 *    the ABI. la_last_error() returns the text of the calling thread's
 *    last error;
 *  - handles own their device memory; host arrays are only read during the
 *    call. One handle must not be used from two threads at once (the
 *    reference's functors have the same rule: their const operator()
 *    mutates scratch, sparse_mat_op/spmv.hpp:181-184).
 *
 * Reference interfaces each entry point replaces are cited file:line.
 */
#ifndef LAAMD_H
#define LAAMD_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  LA_OK = 0,
  LA_ERR_ARG = 1,          /* bad argument (sizes, null pointers, enum)   */
  LA_ERR_HIP = 2,          /* HIP runtime error                           */
  LA_ERR_ALLOC = 3,        /* device or host allocation failed            */
  LA_ERR_NOT_CONVERGED = 4,/* iterative solve hit its iteration limit     */
  LA_ERR_NUMERIC = 5,      /* missing diagonal / negative or zero pivot   */
  LA_ERR_STATE = 6,        /* handle not prepared for this call           */
  LA_ERR_TIMEOUT = 7       /* a sync-free wait exceeded its spin bound    */
} la_status;

typedef enum { LA_F64 = 0, LA_F32 = 1, LA_BF16 = 2 } la_dtype;

/* sparse_mat_op/spmv.hpp: SerialSPMV/ParallelSPMV (:12-57) -> ROW,
 * SegSumSPMV (:59-185) -> SEGSUM, ALBUSSPMV (:186-300) -> ALBUS. AUTO picks
 * SELL when rows are near-uniform (little padding), else ALBUS. */
typedef enum {
  LA_SPMV_ROW = 0,
  LA_SPMV_SEGSUM = 1,
  LA_SPMV_ALBUS = 2,
  LA_SPMV_AUTO = 3,
  LA_SPMV_SELL = 4, /* sliced ELL copy, one lane per row (no counterpart in the
                       reference; AUTO picks it for near-uniform row lengths) */
  LA_SPMV_COLBLOCK = 5 /* row panels x column blocks: each workgroup keeps a
                          panel of rows' sums in LDS while it sweeps blocks of
                          65,536 columns (512 KB of FP64 x, held by the L2
                          caches) in order, and writes y once (no counterpart
                          in the reference; AUTO picks it for ragged rows
                          whose columns are scattered over an x larger than
                          the L2 caches) */
} la_spmv_alg;

/* triangle_solve.hpp:147-208 OptimizedTriangularSolve modes: Barrier ->
 * LEVEL (one launch per level), NoBarrier -> SYNCFREE (one launch, per-row
 * readiness). JACOBI has no counterpart in the reference: k Jacobi sweeps
 * on the triangle (k set per plan, default 3), exact for k >= levels - 1,
 * otherwise an approximate solve for use as a preconditioner. */
typedef enum { LA_TRSV_LEVEL = 0, LA_TRSV_SYNCFREE = 1, LA_TRSV_JACOBI = 2 } la_trsv_mode;

typedef enum { LA_LOWER = 0, LA_UPPER = 1 } la_uplo;

const char *la_last_error(void);
int la_version(void);

/* Device buffers for callers that hold host arrays (the C++ drop-in layer
 * in include/laamd/, whose reference signatures take host pointers). No
 * reference counterpart: the reference has no device. */
int la_device_alloc(size_t bytes, void **ptr);
int la_device_free(void *ptr);
int la_memcpy_h2d(void *dst, const void *src, size_t bytes);
int la_memcpy_d2h(void *dst, const void *src, size_t bytes);
int la_device_synchronize(void);

/* ====================== host-side algorithms (no GPU) ================== */

/* matrix_utils.hpp:637-675 TopologicalSort2<L|U>. iperm[n] (base-adjusted),
 * prefix[nlevels+1] (caller gives n+1 ints). Input strictly triangular. */
int la_topo_sort2(int n, int base, const int *ai, const int *aj, int upper,
                  int *iperm, int *prefix, int *nlevels);
/* matrix_utils.hpp:720-803 SplitLDU. Two-phase: pass NULL Laj/Lav/Uaj/Uav
 * to get Lai, Uai, *lnnz, *unnz; then call again with buffers. */
int la_split_ldu(int n, int base, const int *ai, const int *aj,
                 const double *av, int *Lai, int *Laj, double *Lav, double *D,
                 int *Uai, int *Uaj, double *Uav, int *lnnz, int *unnz);
/* matrix_utils.hpp:805-879 SplitTriangle<U|L> (keeps the diagonal). */
int la_split_triangle(int n, int base, const int *ai, const int *aj,
                      const double *av, int upper, int *Tai, int *Taj,
                      double *Tav, int *tnnz);
/* matrix_utils.hpp:677-704 Diagonal. *all_present = 0 if one is missing. */
int la_diagonal(int n, int base, const int *ai, const int *aj,
                const double *av, int *diagpos, double *diag,
                int *all_present);
/* matrix_utils.hpp:311-390 ParallelTranspose2 (stable). tv may be NULL. */
int la_transpose(int rows, int cols, int base, const int *ai, const int *aj,
                 const double *av, int *ti, int *tj, double *tv);
/* matrix_utils.hpp:429-472 permute: rows by iperm, columns by perm (either
 * may be NULL). */
int la_permute(int n, int base, const int *ai, const int *aj,
               const double *av, const int *iperm, const int *perm, int *pai,
               int *paj, double *pav);
/* reordering/Reordering.cpp:37-136 SerialCM (reverse Cuthill-McKee).
 * iperm[new] = old + base; perm its inverse. Runs with a team of host
 * threads (LAAMD_HOST_THREADS, else the hardware's, at most 16) on
 * structurally symmetric patterns, serially otherwise; the result is the
 * serial definition's, bit for bit, either way. */
int la_rcm(int n, int base, const int *ai, const int *aj, int *iperm,
           int *perm);
/* la_rcm with an explicit thread count (<= 0: the default); the same
 * result for every count. */
int la_rcm_threads(int n, int base, const int *ai, const int *aj, int nthreads,
                   int *iperm, int *perm);
/* mkl_wrapper/incomplete_lu.cpp:153-275 incomplete_lu_k::symbolic_factorize.
 * Two-phase (faj NULL: fai and *nnz only). first_upper[n] are 0-based
 * positions of the first column > i. */
int la_iluk_symbolic(int n, int base, const int *ai, const int *aj, int level,
                     int *fai, int *faj, int *first_upper, int64_t *nnz);
/* mkl_wrapper/incomplete_lu.cpp:277-325 numeric_factorize: combined LU
 * storage (strict-L multipliers, diagonal, strict U). */
int la_iluk_numeric(int n, int base, const int *ai, const int *aj,
                    const double *av, const int *fai, const int *faj,
                    const int *first_upper, double *fav);
/* sparse_mat_op/precond.hpp:39-143 ICCLevelSymbolic (upper triangle input
 * incl. diagonal, diag_pos base-adjusted). Two-phase. */
int la_icc_symbolic(int n, int base, const int *ai, const int *aj,
                    const int *diag_pos, int lvl, int *icc_ai, int *icc_aj,
                    int64_t *nnz);
/* sparse_mat_op/precond.hpp:615-693 ICCLevelNumeric with multiplicative
 * diagonal shift omega; *ok = 0 on a negative pivot. */
int la_icc_numeric(int n, int base, const int *ai, const int *aj,
                   const double *av, const int *diag_pos, double omega,
                   const int *icc_ai, const int *icc_aj, double *icc_av,
                   int *ok);
/* mkl_wrapper/incomplete_cholesky.cpp:195-274 incomplete_cholesky_k::
 * numeric_factorize, one pass: same pattern and arguments as la_icc_numeric,
 * but the diagonal gets the ADDITIVE shift, rows are visited in the
 * reference's first-in first-out column order, and *ok = 0 on a pivot <= 0.
 * The caller runs the retry loop: shift_0 = 1e-3 - min_i a_ii when
 * min_i a_ii <= 0, else 0; then shift = max(1e-3, 2 shift), at most 20
 * retries. */
int la_ick_numeric(int n, int base, const int *ai, const int *aj,
                   const double *av, const int *diag_pos, double shift,
                   const int *icc_ai, const int *icc_aj, double *icc_av,
                   int *ok);

/* ====================== seeded synthetic inputs (host) ================= */
/* Two-phase where the nnz is not closed-form: pass aj == NULL to get ai and
 * *nnz, then call again with aj/av. Values are deterministic for a seed. */
int la_gen_poisson2d(int nx, int ny, int *ai, int *aj, double *av,
                     int64_t *nnz);
int la_gen_laplace3d(int nx, int ny, int nz, int *ai, int *aj, double *av,
                     int64_t *nnz);
int la_gen_powerlaw(int n, double mean_len, double sigma, int max_len,
                    uint64_t seed, int *ai, int *aj, double *av,
                    int64_t *nnz);
int la_gen_aniso27(int nx, int ny, int nz, double aniso, uint64_t seed,
                   int *ai, int *aj, double *av, int64_t *nnz);
/* 2-D 5-point variable-coefficient diffusion: ne = 2*nx*ny-nx-ny interior
 * edges; theta[ne]; A_ii = sum of incident theta + 1 per missing neighbour
 * (Dirichlet), A_ij = -theta_e. edge_i/edge_j[ne] list the edge ends;
 * val_map[4*ne] gives, per edge, the CSR positions of (i,j),(j,i),(i,i),
 * (j,j). */
int la_gen_diffusion2d(int nx, int ny, const double *theta, int *ai, int *aj,
                       double *av, int *edge_i, int *edge_j, int *val_map,
                       int64_t *nnz);
int la_gen_uniform(int64_t n, uint64_t seed, double lo, double hi,
                   double *out);
int la_gen_lognormal(int64_t n, uint64_t seed, double mu, double sigma,
                     double *out);

/* ====================== device CSR matrix ============================== */
typedef struct la_mat_s *la_mat_t;

/* Upload a CSR matrix (host arrays, any base) for SpMV: the handle behind
 * SPMV<CSR,Impl>::setMatrix (sparse_mat_op/spmv.hpp:317) and
 * mkl_sparse_mat (mkl_wrapper/mkl_sparse_mat.h:9-123). Values kept in FP64;
 * la_mat_add_dtype adds an FP32 or BF16 (round-to-nearest-even) copy. */
int la_mat_create(int rows, int cols, int base, const int *ai, const int *aj,
                  const double *av, la_mat_t *out);
int la_mat_destroy(la_mat_t A);
int la_mat_add_dtype(la_mat_t A, la_dtype dt);
/* new values on the same pattern (host or device FP64 array, nnz long);
 * every stored precision is refreshed */
int la_mat_set_values(la_mat_t A, const double *av_host);
int la_mat_set_values_device(la_mat_t A, const double *av_dev);
int la_mat_set_stream(la_mat_t A, void *hip_stream);
int la_mat_info(la_mat_t A, int *rows, int *cols, int64_t *nnz);
/* SegSumSPMV/ALBUSSPMV::preprocess (sparse_mat_op/spmv.hpp:68-77,195-228):
 * nnz-balanced segment table. Called implicitly on first use. */
int la_spmv_preprocess(la_mat_t A, la_spmv_alg alg);
/* The kernel family a request runs on for this matrix: AUTO and SELL give
 * SELL when the sliced-ELL padding stays small; otherwise AUTO gives
 * COLBLOCK for scattered columns over a large x, else ALBUS; SELL falls back
 * to ALBUS; SEGSUM runs on ALBUS's kernels; an empty matrix runs ROW.
 * Builds what the choice needs.
 * No counterpart in the reference (its algorithm is a template argument). */
int la_spmv_resolve(la_mat_t A, la_spmv_alg alg, int *chosen);
/* The sliced-ELL copy (built on demand): *built 1 if the matrix has one,
 * *col16 1 if its columns are 16-bit offsets from the row, *slots the
 * stored slots (nnz plus padding). No counterpart in the reference. */
int la_mat_sell_info(la_mat_t A, int *built, int *col16, int64_t *slots);
/* y = alpha*op(A)*x + beta*y (device pointers), the functor call
 * operator()(size, base, ai, aj, av, b, x, alpha, beta) of
 * sparse_mat_op/spmv.hpp:16-29 with b -> x and x -> y, and
 * mkl_sparse_mat::mult_vec (mkl_sparse_mat.cpp:325-329). beta == 0 never
 * reads y (the SerialSPMV rule, spmv.hpp:26). val: stored precision to use;
 * vec: precision of x and y; acc: accumulation (LA_F64 only with FP64
 * values and vectors). */
int la_spmv(la_mat_t A, la_spmv_alg alg, la_dtype val, la_dtype vec,
            la_dtype acc, double alpha, const void *x, double beta, void *y);
/* y = alpha*A*x + beta*z with a separate input z (residual form b - A x). */
int la_spmv_axpby(la_mat_t A, la_spmv_alg alg, la_dtype val, la_dtype vec,
                  la_dtype acc, double alpha, const void *x, double beta,
                  const void *z, void *y);
/* y = alpha*A^T*x + beta*y: mkl_sparse_mat::transpose_mult_vec
 * (mkl_sparse_mat.cpp:331-336); builds A^T on first use. */
int la_spmv_transpose(la_mat_t A, la_dtype val, la_dtype vec, double alpha,
                      const void *x, double beta, void *y);
/* gradient of a loss w.r.t. the stored values: g[k] = -lambda[row k] *
 * x[col k] for every stored entry (FP64, device). */
int la_csr_value_grad(la_mat_t A, const double *lambda, const double *x,
                      double *g);

/* ====================== sparse triangular solve ======================== */
typedef struct la_trsv_s *la_trsv_t;

/* OptimizedTriangularSolve::analysis (triangle_solve.tpp:17-113) and the
 * level sets of TopologicalSort2: (ai,aj,av) is a STRICT triangle, diag the
 * diagonal (NULL = unit). Lower solves sweep forward
 * (triangle_solve.hpp:32-43), upper backward (:59-71):
 * x[i] = (b[i] - sum_j T_ij x_j) / diag[i]. */
int la_trsv_create(int n, int base, const int *ai, const int *aj,
                   const double *av, const double *diag, la_uplo uplo,
                   la_trsv_t *out);
int la_trsv_destroy(la_trsv_t T);
int la_trsv_add_dtype(la_trsv_t T, la_dtype dt);
int la_trsv_set_stream(la_trsv_t T, void *hip_stream);
int la_trsv_info(la_trsv_t T, int *n, int *nlevels, int64_t *nnz);
/* Which schedule the sync-free mode runs (no counterpart in the reference):
 * *tiled = 1 when the triangle's pattern was recognised as a grid (from the
 * pattern, or for a natural-order box grid from the row indices) and cut
 * into *tiles tiles of edge *tile_edge grid steps, one workgroup each (the
 * environment variable LAAMD_SF_TILED=0 at plan creation turns it off);
 * 0 for the level-chunk schedule. */
int la_trsv_schedule_info(la_trsv_t T, int *tiled, int *tiles, int *tile_edge);
/* Sweep count of LA_TRSV_JACOBI solves with this plan (0 = diagonal scaling
 * only); no counterpart in the reference. */
int la_trsv_set_sweeps(la_trsv_t T, int sweeps);
/* OptimizedTriangularSolve::operator()(b, x) (triangle_solve.tpp:115-125);
 * b and x are distinct device vectors. */
int la_trsv_solve(la_trsv_t T, la_trsv_mode mode, la_dtype val, la_dtype vec,
                  const void *b, void *x);
/* (T + D)^T x = b, ForwardSubstitutionT's problem (triangle_solve.hpp:
 * 87-98) for the adjoint; builds the transposed plan on first use. */
int la_trsv_solve_transpose(la_trsv_t T, la_trsv_mode mode, la_dtype val,
                            la_dtype vec, const void *b, void *x);

/* ====================== ILU(k) / IC(k) apply =========================== */
typedef struct la_prec_s *la_prec_t;

/* incomplete_lu_base::solve (mkl_wrapper/incomplete_lu.cpp:36-48): factor
 * in combined LU storage (ILU(k) numeric output). */
int la_prec_create_ilu(int n, int base, const int *fai, const int *faj,
                       const double *fav, la_prec_t *out);
/* incomplete_choleksy_base::solve (mkl_wrapper/incomplete_cholesky.cpp:
 * 21-29): upper factor R incl. diagonal, M = R^T R. */
int la_prec_create_ic(int n, int base, const int *rai, const int *raj,
                      const double *rav, la_prec_t *out);
/* Preconditioners whose numeric factorisation runs on the device (no
 * counterpart in the reference, whose numeric phase is serial host code:
 * incomplete_lu.cpp:277-325, precond.hpp:615-693). Creation runs the host
 * symbolic phase (iluk_symbolic / ICCLevelSymbolic on the given CSR of A;
 * IC takes the full symmetric A and uses its upper triangle) and builds the
 * solve plans; la_prec_refactor(P, a_vals) factors on the device from A's
 * values (device FP64 array in that CSR's order) with the reference's
 * arithmetic and order -- bitwise the host factor -- and may be called
 * again with new values. LA_ERR_NUMERIC: missing diagonal (ILU) or negative
 * pivot (IC). la_prec_factor copies the factor to the host: the pattern
 * 0-based (fai: n+1, faj/fav: *nnz entries); pass null arrays to get *nnz. */
int la_prec_create_ilu_gpu(int n, int base, const int *ai, const int *aj, int level,
                           la_prec_t *out);
int la_prec_create_ic_gpu(int n, int base, const int *ai, const int *aj, int level, double omega,
                          la_prec_t *out);
/* The same with incomplete_cholesky_k's numeric arithmetic
 * (mkl_wrapper/incomplete_cholesky.cpp:195-274, one pass of its retry
 * loop): `shift` is ADDED to the diagonal, rows are eliminated in that
 * class's first-in first-out column order, and a pivot <= 0 gives
 * LA_ERR_NUMERIC; the factor is bitwise la_ick_numeric's. */
int la_prec_create_ick_gpu(int n, int base, const int *ai, const int *aj, int level,
                           double shift, la_prec_t *out);
int la_prec_refactor(la_prec_t P, const double *a_vals);
int la_prec_factor(la_prec_t P, int *fai, int *faj, double *fav, int64_t *nnz);
int la_prec_destroy(la_prec_t P);
int la_prec_add_dtype(la_prec_t P, la_dtype dt);
int la_prec_set_stream(la_prec_t P, void *hip_stream);
/* z = M^-1 r (device vectors). */
int la_prec_apply(la_prec_t P, la_trsv_mode mode, la_dtype val, la_dtype vec,
                  const void *r, void *z);
/* z = M^-T r. */
int la_prec_apply_transpose(la_prec_t P, la_trsv_mode mode, la_dtype val,
                            la_dtype vec, const void *r, void *z);
int la_prec_info(la_prec_t P, int *levels_lower, int *levels_upper);
/* the tile count of each triangle's tiled sync-free schedule, 0 where the
 * level-chunk schedule runs (no counterpart in the reference) */
int la_prec_schedule_info(la_prec_t P, int *tiled_lower, int *tiled_upper);
/* Sweep count of LA_TRSV_JACOBI applies, both triangles (default 3); no
 * counterpart in the reference. */
int la_prec_set_sweeps(la_prec_t P, int sweeps);

/* ====================== fused BLAS-1 (FP64 accumulation) =============== */
int la_dot(int64_t n, la_dtype vec, const void *x, const void *y,
           double *result_host, void *hip_stream);
/* la_dot without any allocation, for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc);
 * no counterpart in the reference. The result goes to device memory
 * (*result_dev) on the given stream and the call returns without waiting.
 * The partial sums live in a fixed per-GPU device array of 512 doubles
 * (4 KiB, held until the process ends), whatever n is; calls that overlap
 * on one GPU share it and race. Bitwise equal to la_dot. */
int la_dot_noalloc(int64_t n, la_dtype vec, const void *x, const void *y,
                   double *result_dev, void *hip_stream);
/* y += a*x; *result_host = <y, z> (z may be y) */
int la_axpy_dot(int64_t n, la_dtype vec, double a, const void *x, void *y,
                const void *z, double *result_host, void *hip_stream);

/* ====================== Krylov solvers ================================= */
typedef enum {
  LA_PREC_FP64 = 0,       /* everything FP64                              */
  LA_PREC_MIXED_F32 = 1,  /* inner solve: FP32 matrix/factors/vectors      */
  LA_PREC_MIXED_BF16 = 2, /* inner solve: BF16 matrix+factors, FP32 vectors*/
  LA_PREC_MIXED_BF16_F32P = 3, /* inner solve: BF16 matrix, FP32 factors and
                                  vectors (no counterpart in the reference)  */
  LA_PREC_MIXED_F32_BF16P = 4   /* inner solve: FP32 matrix and vectors, BF16
                                  factors (no counterpart in the reference)  */
} la_precision;

typedef enum {
  LA_STOP_TRUE_RESIDUAL = 0, /* ||b-Ax||/||b|| every iteration, with an
                                extra SpMV (mkl_solver.cpp:211-225)        */
  LA_STOP_RECURRENCE = 1     /* recurrence residual, confirmed by a true
                                residual before returning                  */
} la_stop_rule;

typedef struct {
  double rel_tol;      /* mkl_iterative_solver::set_rel_tol (default 1e-8) */
  double abs_tol;      /* set_abs_tol (default 0)                          */
  int max_iters;       /* set_max_iters (default 1000)                     */
  int restart;         /* FGMRES: set_restart_steps (default 30; 0 runs
                          unrestarted GMRES up to max_iters, no upper cap) */
  int stop_rule;       /* la_stop_rule                                      */
  int precision;       /* la_precision                                     */
  int trsv_mode;       /* la_trsv_mode for the preconditioner apply (default SYNCFREE) */
  int spmv_alg;        /* la_spmv_alg                                      */
  double inner_tol;    /* mixed: inner relative reduction per refinement (the deepest asked; a
                          step stops earlier once that reaches rel_tol/abs_tol, margin 4) */
  int max_outer;       /* mixed: refinement steps                          */
  int print_level;     /* set_print_level: 1 prints each residual          */
  double reliable_delta; /* mixed PCG: > 0 (default 0.1) runs reliable
                          updates -- one FP32 iteration whose residual is
                          replaced by the FP64 true residual (and x updated
                          in FP64) when it has fallen by this factor; <= 0
                          runs restarted refinement (inner_tol, max_outer).
                          No counterpart in the reference.                  */
} la_solver_opts;

typedef struct {
  int converged;
  int iters;           /* total (inner) iterations                         */
  int outer_iters;     /* refinement steps (mixed), 1 otherwise            */
  double rel_residual; /* final true ||b-Ax||/||b|| in FP64                */
  double seconds;      /* wall time of the device-resident loop            */
} la_solver_result;

void la_solver_opts_default(la_solver_opts *o);
/* mkl_pcg_solver::solve (mkl_wrapper/mkl_solver.cpp:150-251). b, x device
 * FP64 vectors; x holds x0 on entry. P may be NULL. */
int la_pcg_solve(la_mat_t A, la_prec_t P, const la_solver_opts *opts,
                 const double *b, double *x, la_solver_result *res);
/* mkl_fgmres_solver::solve (mkl_wrapper/mkl_solver.cpp:253-386): right
 * (flexible) preconditioned GMRES(restart). */
int la_fgmres_solve(la_mat_t A, la_prec_t P, const la_solver_opts *opts,
                    const double *b, double *x, la_solver_result *res);
/* A^T x = b with the same loops, the transposed product
 * (mkl_sparse_mat::transpose_mult_vec, mkl_sparse_mat.cpp:331-336) and the
 * transposed preconditioner apply M^-T; the adjoint solve of x = A^-1 b. */
/* FGMRES with a right operator R (mkl_fgmres_solver(A, P, R),
 * mkl_wrapper/mkl_solver.cpp:319-324): the Krylov operator is A R^-1 (R
 * applied as a preconditioner, then A) and P the flexible preconditioner;
 * x returns u with A R^-1 u = b, as the reference does (the solution of
 * A x = b is R^-1 u). FP64 precision only. */
int la_fgmres_solve_right(la_mat_t A, la_prec_t P, la_prec_t R, const la_solver_opts *opts,
                          const double *b, double *x, la_solver_result *res);
int la_pcg_solve_transpose(la_mat_t A, la_prec_t P, const la_solver_opts *opts,
                           const double *b, double *x, la_solver_result *res);
int la_fgmres_solve_transpose(la_mat_t A, la_prec_t P, const la_solver_opts *opts,
                              const double *b, double *x, la_solver_result *res);

/* ====================== symmetric scaling ============================= */
/* rowwiseSqrtNorm (mkl_wrapper/mkl_sparse_mat.cpp:517-533): d_i =
 * 1/sqrt(sqrt(sum_j a_ij^2)) (1 when the sum is not above DBL_MIN);
 * DtAD (:535-544): a_ij *= d_i * d_j, in place. Host arrays: la_*; a device
 * matrix and a device d: la_mat_* (bitwise the host results; la_mat_dtad
 * rescales the matrix's FP32/BF16 copies too). */
int la_rowwise_sqrt_norm(int n, int base, const int *ai, const double *av, double *d);
int la_dtad(int n, int base, const int *ai, const int *aj, double *av, const double *d);
int la_mat_rowwise_sqrt_norm(la_mat_t A, double *d);
int la_mat_dtad(la_mat_t A, const double *d);

/* ====================== readers ======================================= */
/* Matrix files into a host CSR sorted by (row, column):
 *  la_read_matrix_market: coordinate real/integer/pattern, general/
 *    symmetric/skew-symmetric; symmetric files expanded to both triangles
 *    (read_matrix_market_csr, utils/utils.h:93-132);
 *  la_read_binary_coo: the reference's int64-count + 16-byte-record COO
 *    format, indices in `base` (ReadFromBinaryCSR, utils/utils.h:24-91).
 * Query sizes with la_csr_file_info, copy out with la_csr_file_copy (ai:
 * rows+1, aj/av: nnz, in the requested base), release with
 * la_csr_file_free. la_read_binary_vector (ReadFromBinaryVec,
 * utils/utils.cpp:36-46): out null -> *n = length; else *n = capacity in,
 * length out. */
typedef struct la_csr_file_s *la_csr_file_t;
int la_read_matrix_market(const char *path, int base, la_csr_file_t *out);
int la_read_binary_coo(const char *path, int base, la_csr_file_t *out);
int la_csr_file_info(la_csr_file_t F, int *rows, int *cols, int64_t *nnz);
int la_csr_file_copy(la_csr_file_t F, int *ai, int *aj, double *av);
int la_csr_file_free(la_csr_file_t F);
int la_read_binary_vector(const char *path, double *out, int64_t *n);

#ifdef __cplusplus
}
#endif
#endif
