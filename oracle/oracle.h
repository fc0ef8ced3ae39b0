/*
 * CPU oracle for the sparse hot path: a plain-C FP64 restatement of the
 * reference algorithms (miaodi/linearAlgebra). TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the result to compare against (or the
 * timed CPU baseline), never as the thing measured or shipped. The product
 * path (linearalgebra_amd) never links or calls it.
 *
 * Built with -O2 -ffp-contract=off: the reference's Release build has no
 * -march flag, so x86-64 baseline code carries no FMA and every product is
 * rounded before it is added (CMakeLists.txt:13-14).
 *
 * Index type int32, value type double, base 0 or 1 passed explicitly as in
 * the reference's free functions (sparse_mat_op/spmv.hpp:16-19).
 *
 * Parity status of each function is stated next to it:
 *   pinned   - checked in tests/ against a closed form, a known-answer test
 *              from the reference's own tests, scipy/numpy, or brute force.
 *   unpinned - no independent check exists (see the comment for why).
 */
#ifndef LAAMD_ORACLE_H
#define LAAMD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- SpMV: x = alpha*A*b + beta*x (b is the input, x the output) ---- */
/* sparse_mat_op/spmv.hpp:12-29. pinned (KAT, scipy, dense brute force). */
void orc_spmv_serial(int n, int base, const int *ai, const int *aj,
                     const double *av, const double *b, double *x,
                     double alpha, double beta);
/* sparse_mat_op/spmv.hpp:31-57 + utils/utils.h:188-222. pinned (== serial
 * bitwise, the reference's own 1e-14 test tightened to equality). */
void orc_spmv_parallel(int n, int base, const int *ai, const int *aj,
                       const double *av, const double *b, double *x,
                       double alpha, double beta, int nthreads);
/* sparse_mat_op/spmv.hpp:59-185. pinned (vs serial within re-association). */
void orc_spmv_segsum(int n, int base, const int *ai, const int *aj,
                     const double *av, const double *b, double *x,
                     double alpha, double beta, int nthreads);
/* sparse_mat_op/spmv.hpp:186-300. pinned (vs serial within re-association). */
void orc_spmv_albus(int n, int base, const int *ai, const int *aj,
                    const double *av, const double *b, double *x,
                    double alpha, double beta, int nthreads);

/* ---- triangular solves ---- */
/* triangle_solve.hpp:32-43 (strict L, unit diagonal). pinned. */
void orc_forward_sub(int n, int base, const int *ai, const int *aj,
                     const double *av, const double *b, double *x);
/* triangle_solve.hpp:59-71 (strict U + diag). pinned. */
void orc_backward_sub(int n, int base, const int *ai, const int *aj,
                      const double *av, const double *diag, const double *b,
                      double *x);
/* No counterpart in the reference (the device's JACOBI triangular-solve
 * mode): k Jacobi sweeps, y_0 = D^-1 b, y_{s+1} = D^-1 (b - T y_s), in the
 * column-scaled form w = D y the device runs.
 * pinned (k >= levels - 1 reproduces orc_forward_sub / orc_backward_sub:
 * D^-1 T is nilpotent; k = 0 is D^-1 b; scipy dense Neumann series). */
void orc_trsv_jacobi(int n, int base, const int *ai, const int *aj,
                     const double *av, const double *diag, const double *b,
                     int k, double *x);
/* triangle_solve.hpp:87-98 ((I+U^T) x = b from CSR of strict U). pinned. */
void orc_forward_sub_t(int n, int base, const int *ai, const int *aj,
                       const double *av, const double *b, double *x);
/* matrix_utils.hpp:637-675. upper!=0 sweeps rows descending. iperm[n],
 * prefix[levels+1] (caller provides n+1). Returns #levels. pinned
 * (grid closed form, brute-force longest path). */
int orc_topo_sort2(int n, int base, const int *ai, const int *aj, int upper,
                   int *iperm, int *prefix);
/* triangle_solve.hpp:100-145. pinned (== serial bitwise). */
void orc_level_forward(const int *iperm, const int *prefix, int lvls, int n,
                       int base, const int *ai, const int *aj,
                       const double *av, const double *b, double *x);
void orc_level_backward(const int *iperm, const int *prefix, int lvls, int n,
                        int base, const int *ai, const int *aj,
                        const double *av, const double *diag, const double *b,
                        double *x);

/* ---- CSR helpers ---- */
/* matrix_utils.hpp:720-803. Counts first (Lai/Uai filled, returns via
 * lnnz and unnz), then fill with Laj.. non-NULL. pinned (reconstruction). */
void orc_split_ldu(int n, int base, const int *ai, const int *aj,
                   const double *av, int *Lai, int *Laj, double *Lav,
                   double *D, int *Uai, int *Uaj, double *Uav, int *lnnz,
                   int *unnz);
/* matrix_utils.hpp:805-879; upper!=0 keeps j>=i else j<=i. Two-phase like
 * orc_split_ldu. pinned. */
void orc_split_triangle(int n, int base, const int *ai, const int *aj,
                        const double *av, int upper, int *Tai, int *Taj,
                        double *Tav, int *tnnz);
/* matrix_utils.hpp:677-704. Returns 1 if every diagonal exists. pinned. */
int orc_diagonal(int n, int base, const int *ai, const int *aj,
                 const double *av, int *diagpos, double *diag);
/* matrix_utils.hpp:311-390 (result is the stable transpose). pinned. */
void orc_transpose(int rows, int cols, int base, const int *ai, const int *aj,
                   const double *av, int *ti, int *tj, double *tv);
/* matrix_utils.hpp:429-472 (rows by iperm, columns by perm, insertion
 * sort per row). pinned (round trip). */
void orc_permute(int n, int base, const int *ai, const int *aj,
                 const double *av, const int *iperm, const int *perm,
                 int *pai, int *paj, double *pav);

/* ---- ILU(k) (mkl_wrapper/incomplete_lu.cpp) ---- */
/* :153-275. Writes fai[n+1]; returns nnz of the factor pattern; when faj is
 * non-NULL also writes the pattern (same base as A) and first_upper[n]
 * (0-based position of the first column > i). Call twice (count, fill).
 * pinned (dense brute-force level-of-fill, ILU(0)=pattern(A)). */
int orc_iluk_symbolic(int n, int base, const int *ai, const int *aj,
                      int level, int *fai, int *faj, int *first_upper);
/* :277-325, combined LU storage (strict-L multipliers, diag, strict-U),
 * values in the pattern's base (the reference then shifts to 1-based; the
 * values are unchanged). Returns 0 if a row has no diagonal. pinned (dense
 * IKJ restricted to the pattern; (LU)_ij = a_ij on the pattern). */
int orc_iluk_numeric(int n, int base, const int *ai, const int *aj,
                     const double *av, const int *fai, const int *faj,
                     const int *first_upper, double *fav);

/* ---- IC(k) (sparse_mat_op/precond.hpp) ---- */
/* ICCLevelSymbolic0, :39-143. Input is the UPPER triangle incl. diagonal;
 * diag_pos[j] = position (base-adjusted) of a_jj. Two-phase (count: icc_aj
 * NULL). pinned (== upper part of the ILU(k) pattern of the symmetric
 * matrix, brute force). */
int orc_icc_symbolic(int n, int base, const int *ai, const int *aj,
                     const int *diag_pos, int lvl, int *icc_ai, int *icc_aj);
/* ICCLevelNumeric, :615-693. Returns 1 on success, 0 on negative pivot.
 * pinned (tridiagonal SPD -> exact Cholesky; (U^T U)_ij = a_ij on pattern
 * for IC(0)). */
int orc_icc_numeric(int n, int base, const int *ai, const int *aj,
                    const double *av, const int *diag_pos, double omega,
                    const int *icc_ai, const int *icc_aj, double *icc_av);

/* incomplete_cholesky_k::numeric_factorize, one pass
 * (mkl_wrapper/incomplete_cholesky.cpp:195-274): same pattern and diag_pos
 * as orc_icc_numeric, ADDITIVE diagonal shift, rows visited first-in
 * first-out per column, fails (returns 0) on a pivot <= 0. pinned (the
 * textbook IC(0) recurrence of the 5-point grid in the reference's visit
 * order, bitwise; tridiagonal -> Cholesky of A + shift I; (R^T R)_ij =
 * a_ij + shift d_ij on the IC(0) pattern; a zero pivot fails here and
 * passes orc_icc_numeric; within rounding of orc_icc_numeric, as
 * tests/precond_test.cpp:118-164 prints). */
int orc_ick_numeric(int n, int base, const int *ai, const int *aj,
                    const double *av, const int *diag_pos, double shift,
                    const int *icc_ai, const int *icc_aj, double *icc_av);

/* ---- preconditioned Krylov (mkl_wrapper/mkl_solver.cpp) ---- */
/* prec_kind: 0 none, 1 ILU (combined LU factor CSR, apply trsv(L unit) then
 * trsv(U non-unit), incomplete_lu.cpp:36-48), 2 IC (upper factor incl. diag,
 * apply U^-1 U^-T, incomplete_cholesky.cpp:21-29). The factor uses `base`
 * like A. */
typedef struct {
  int kind;
  const int *fai, *faj;
  const double *fav;
} orc_prec;

/* PCG with the reference's stopping rule: true residual ||b-Ax||/||b|| <
 * rel_tol or ||b-Ax|| < abs_tol checked every iteration (:170,:211-225).
 * Textbook PCG (Saad Alg. 9.1) in place of MKL dcg, whose arithmetic is
 * closed: parity unpinned at the Krylov boundary; pinned against scipy cg
 * solutions, the k-distinct-eigenvalue bound and M=A -> one iteration.
 * Returns 1 on convergence, 0 on max-iteration failure. */
int orc_pcg(int n, int base, const int *ai, const int *aj, const double *av,
            const orc_prec *P, const double *b, double *x, double rel_tol,
            double abs_tol, int maxiter, int *iters, double *relres);
/* Right-preconditioned flexible GMRES(m) with modified Gram-Schmidt and
 * Givens rotations (Saad Alg. 9.6) in place of MKL dfgmres; stops on the
 * Arnoldi residual estimate / ||b|| (:335-346) or the Hessenberg breakdown
 * h_{k+1,k} < 1e-16 (:353-367). Same parity status as orc_pcg. */
int orc_fgmres(int n, int base, const int *ai, const int *aj,
               const double *av, const orc_prec *P, const double *b,
               double *x, int restart, double rel_tol, double abs_tol,
               int maxiter, int *iters, double *relres);
/* z = M^-1 r for the preconditioners above. */
void orc_prec_apply(int n, int base, const orc_prec *P, const double *r,
                    double *z);

/* ---- symmetric scaling (mkl_wrapper/mkl_sparse_mat.cpp:517-544) ---- */
/* rowwiseSqrtNorm: d_i = 1/sqrt(sqrt(sum_j a_ij^2)) (1 if the sum is not
 * above DBL_MIN), summed left to right. Pinned by a Python loop in the same
 * order and by d_i^4 * sum_j a_ij^2 == 1 within rounding. */
void orc_rowwise_sqrt_norm(int n, int base, const int *ai, const double *av,
                           double *d);
/* DtAD: a_ij *= d_i * d_j in place. Pinned by the dense D A D product. */
void orc_dtad(int n, int base, const int *ai, const int *aj, double *av,
              const double *d);

/* ---- RCM (reordering/Reordering.cpp:37-136 and what it calls) ---- */
/* iperm[new] = old + base, perm = inverse. The reference's BFS width member
 * is uninitialised (reordering/BFS.h:47); here it starts at 0 for every BFS
 * object PseudoDiameter constructs and is never reset between the calls on
 * one object, as there. Where the reference would loop forever (no
 * candidate left in the last level, Reordering.h:91-130 with every degree
 * overwritten), the search stops and keeps the current source.
 * pinned: exact iperm of four small graphs derived by hand from
 * Reordering.cpp:37-136, Reordering.h:73-136 and BFS.cpp:10-180 (the
 * degree overwrite decides a child order; a second search round; the
 * never-reset width decides the end point; components of sizes 1, 2, 4
 * and 5), the BFS known-answer test (tests/bfs_test.cpp:12-51),
 * permutation validity, bandwidth reduction, base/run determinism. */
void orc_rcm(int n, int base, const int *ai, const int *aj, int *iperm,
             int *perm);
/* Serial BFS levels (reordering/BFS.cpp:10-57); returns height. */
int orc_bfs_levels(int n, int base, const int *ai, const int *aj, int source,
                   int *levels);

/* Planted mistakes (tests of the tests only; 0 = none, the default). */
enum {
  ORC_PLANT_NONE = 0,
  ORC_PLANT_SPMV_DROP_LAST = 1,       /* SerialSPMV skips each row's last entry */
  ORC_PLANT_RCM_ORIGINAL_DEGREES = 2, /* RCM child sort by the unmodified degrees */
  ORC_PLANT_RCM_WIDTH_RESET = 3       /* PseudoDiameter resets the width per BFS */
};
void orc_set_plant(int which);

/* timing helper for the CPU baseline: seconds of wall clock */
double orc_wtime(void);
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
