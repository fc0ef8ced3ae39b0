// Device-side objects of the hot path (matrix, triangular-solve plan,
// preconditioner) and the launchers the host code calls.
#pragma once
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "common.hpp"

namespace laamd {

// -------------------------------------------------------------- CSR matrix
struct Mat {
  int rows = 0, cols = 0;
  int64_t nnz = 0;
  DevBuf ai, aj;       // 0-based int32 row pointers / columns
  MultiValues vals;    // stored values in up to three precisions
  hipStream_t stream = nullptr;
  double avg_row = 0;  // nnz / rows
  int max_row = 0;
  // nnz-balanced segment table (ALBUS / SEGSUM)
  bool seg_ready = false;
  int64_t nseg = 0;
  DevBuf seg_first_row, seg_head, seg_tail, seg_tail_row, seg_tail_end;
  // sliced ELL copy (slices of 64 rows, column-major inside a slice): built
  // on first use; values gathered from `vals` per precision when stale
  int sell_state = 0;  // 0 not built, 1 built, -1 not worth it (too much padding)
  int64_t sell_nslices = 0, sell_len = 0;
  int sell_uniform_w = 0;  // > 0 when every slice has this width
  bool sell_col16 = false; // columns stored as 16-bit offsets from the row
  DevBuf sell_ptr, sell_w, sell_col, sell_src, sell_val[3];
  uint64_t sell_ver[3] = {0, 0, 0};
  // row panels x column blocks (LA_SPMV_COLBLOCK), built on first use
  // (cb_state: 0 not decided, 1 built, -1 not worth it): pb_np panels of rows
  // (pb_prow), pb_nb blocks of pb_width columns; per (panel, block) tile the
  // chunk offsets of its wavefronts (pb_chunk), the packed entries (pb_idx)
  // and the entry of A each holds (pb_src, -1 for padding); the values
  // follow A's through pb_src (pb_ver: the version gathered)
  int cb_state = 0;
  int pb_np = 0, pb_nb = 0, pb_width = 0;
  int64_t pb_len = 0;
  DevBuf pb_prow, pb_chunk, pb_idx, pb_src;
  MultiValues pb_val;
  uint64_t pb_ver = ~0ull;
  // A^T, its pattern built on first transposed product; its values are
  // gathered on the device from A's (t_src: the entry of A each entry of
  // A^T holds) whenever A's values changed (t_ver: the version gathered)
  std::unique_ptr<Mat> transposed;
  DevBuf t_src;
  uint64_t t_ver = ~0ull;
  int sym_pattern = -1;  // structurally symmetric: -1 not known yet
  // host copy of the pattern, kept for the transpose and the value map
  std::vector<int> h_ai, h_aj;
};

std::unique_ptr<Mat> make_mat(int rows, int cols, int base, const int *ai, const int *aj,
                              const double *av, hipStream_t s);
void spmv_prepare_segments(Mat &A);
// builds the sliced-ELL copy; false when the matrix is not suited to it
bool spmv_prepare_sell(Mat &A);
// y = rs o (z - A x) with the sliced-ELL kernel (rs: row scale in the value
// precision); false, with nothing launched, when the matrix has no
// sliced-ELL form. FP64 accumulation for FP64 values and vectors.
bool spmv_sell_rowscaled(Mat &A, la_dtype val, la_dtype vec, const void *rs, const void *x,
                         const void *z, void *y, hipStream_t s);
// the kernel family a request runs on (SEGSUM shares ALBUS's kernels)
la_spmv_alg spmv_resolve(Mat &M, la_spmv_alg alg);
// rowwiseSqrtNorm / DtAD on the device matrix (d: device, rows entries)
void mat_rowwise_sqrt_norm(Mat &A, double *d, hipStream_t s);
void mat_dtad(Mat &A, const double *d, hipStream_t s);
// y = alpha*A*x + beta*z (z may equal y; beta == 0 never reads z)
void spmv(Mat &A, la_spmv_alg alg, la_dtype val, la_dtype vec, la_dtype acc, double alpha,
          const void *x, double beta, const void *z, void *y, hipStream_t s);
void spmv_transpose(Mat &A, la_dtype val, la_dtype vec, double alpha, const void *x,
                    double beta, void *y, hipStream_t s);
// A^T as its own matrix: the pattern once (stable host transpose), the
// values gathered on the device whenever A's change; lower-precision copies
// follow A's
Mat &ensure_transposed(Mat &A, hipStream_t s);
// A's pattern is structurally symmetric (checked once)
bool mat_symmetric_pattern(Mat &A);
void csr_value_grad(Mat &A, const double *lambda, const double *x, double *g, hipStream_t s);

// -------------------------------------------------------------- SpTRSV plan
struct TrsvPlan {
  int n = 0, nlevels = 0;
  int64_t nnz = 0;
  bool upper = false, unit = true;
  int G = 1;                   // lanes per row
  DevBuf rowptr, col, perm;    // rows in level order; columns keep original ids
  MultiValues val, diag;       // values / reciprocal diagonal in level order
  std::vector<int> level_ptr;  // host, nlevels+1
  int nchunks = 0;
  DevBuf chunk_start, counter, err_flag;
  // sync-free launch shape: confined to one XCD (plain stores, L2-latency
  // polls) or spread over all XCDs; workgroups taking part
  bool sf_xlocal = true;
  unsigned sf_wgs = 32;
  int sf_nt = 256;  // threads per workgroup; a chunk holds sf_nt/G rows
  hipStream_t stream = nullptr;
  // host copy of the strict triangle + diagonal (0-based), for the transpose
  std::vector<int> h_ai, h_aj;
  std::vector<double> h_av, h_diag;
  std::unique_ptr<TrsvPlan> transposed;
  // level order on the host (position k -> row)
  std::vector<int> h_perm;
  // factors computed on the device: where each stored value (original
  // entry order) and each diagonal entry sits in the factor's value array,
  // the level-ordered device copies of the same, and the factor array the
  // values were last gathered from
  std::vector<int64_t> h_src, h_dsrc;
  DevBuf val_src, diag_src;
  const double *fsrc = nullptr;
  // captured per-level launch sequences, keyed by (val, vec, b, x)
  std::map<std::tuple<int, int, const void *, void *>, hipGraphExec_t> graphs;
  // tiled sync-free schedule (built when the pattern is recognised as a
  // grid): one workgroup per tile, its rows' dependencies inside the tile
  // kept in LDS. Rows are stored group by group, 64 slots per group (one
  // per lane; slots past the group's rows are padding): the row id (-1 for
  // padding), t_maxe 16-bit LDS indices of the values each row reads, and
  // per precision its values and reciprocal diagonal (refreshed with the
  // level-ordered values)
  bool tiled = false;
  int t_nparts = 0, t_ngroups = 0, t_mmax = 0, t_hmax = 0, t_maxe = 0, t_tau = 0;
  int t_rlanes = 1;  // lanes per row (kTileRowLanes for rows of more than 8 entries)
  double t_remote = 0;
  DevBuf t_gptr, t_prow, t_pcode, t_rfptr, t_rfcol, t_rfslot;
  DevBuf t_lsrc, t_dpos;  // slot entry -> level-ordered entry; slot -> level position (-1: padding)
  MultiValues t_val, t_dinv;
  unsigned t_wgs = 0;   // workgroups launched
  std::string t_why;    // why the tiled schedule was not built
  // JACOBI mode (built on first use): x = D^-1 w_k with w_0 = b and
  // w_{s+1} = b - (N D^-1) w_s, N the strict triangle in row order
  int sweeps = 3;
  std::unique_ptr<Mat> jac;  // N D^-1
  MultiValues jdinv;         // D^-1 (non-unit triangles)
  DevBuf jtmp[2];            // sweep vectors
  DevBuf jsrc, jdsrc;        // device-factor sources, row order
  ~TrsvPlan();
};

std::unique_ptr<TrsvPlan> make_trsv(int n, int base, const int *ai, const int *aj,
                                    const double *av, const double *diag, bool upper,
                                    hipStream_t s);
void trsv_add_dtype(TrsvPlan &T, la_dtype dt);
void trsv_solve(TrsvPlan &T, la_trsv_mode mode, la_dtype val, la_dtype vec, const void *b,
                void *x, hipStream_t s);
TrsvPlan &trsv_transposed(TrsvPlan &T);
// records where the plan's values live in a device factor array (src per
// entry of the triangle as given to make_trsv, dsrc per row; dsrc may be
// null for a unit diagonal)
void trsv_set_sources(TrsvPlan &T, const int64_t *src, const int64_t *dsrc, hipStream_t s);
// gathers the plan's values (and 1/diagonal) from the factor array f, then
// refreshes the FP32/BF16 copies and any transposed plan
void trsv_refresh(TrsvPlan &T, const double *f, hipStream_t s);
// JACOBI mode sweep count for T and its transposed plan (0..1000000)
void trsv_set_sweeps(TrsvPlan &T, int sweeps);
// checks (and clears) the sync-free spin-bound error flag; synchronizes s
bool trsv_check_timeout(TrsvPlan &T, hipStream_t s);

// -------------------------------------------------------------- ILU / IC
struct NumericPlan;
struct NumericPlanDeleter {
  void operator()(NumericPlan *p) const;
};
struct Prec {
  bool ic = false;
  int n = 0;
  std::unique_ptr<TrsvPlan> lower, upper;
  DevBuf tmp[2];  // intermediate vector, FP64 and FP32
  hipStream_t stream = nullptr;
  // set when the factor is computed on the device (make_prec_*_gpu)
  std::unique_ptr<NumericPlan, NumericPlanDeleter> num;
};
std::unique_ptr<Prec> make_prec_ilu(int n, int base, const int *fai, const int *faj,
                                    const double *fav, hipStream_t s);
std::unique_ptr<Prec> make_prec_ic(int n, int base, const int *rai, const int *raj,
                                   const double *rav, hipStream_t s);
// device numeric factorisation: symbolic phase on the host at creation,
// numeric phase (same operation order as the host's, bitwise the same
// factor) on the device at each prec_refactor
std::unique_ptr<Prec> make_prec_ilu_gpu(int n, int base, const int *ai, const int *aj, int level,
                                        hipStream_t s);
// additive: incomplete_cholesky_k's arithmetic (omega is then the added
// shift, rows first-in first-out, a pivot <= 0 fails)
std::unique_ptr<Prec> make_prec_ic_gpu(int n, int base, const int *ai, const int *aj, int level,
                                       double omega, hipStream_t s, bool additive = false);
// a_vals: device FP64 values of A in the CSR layout given at creation
void prec_refactor(Prec &P, const double *a_vals, hipStream_t s);
// factor values (ILU: combined LU; IC: upper R) and pattern, host copies
void prec_factor(Prec &P, std::vector<int> *fai, std::vector<int> *faj, std::vector<double> *fav,
                 hipStream_t s);
void prec_add_dtype(Prec &P, la_dtype dt);
void prec_apply(Prec &P, la_trsv_mode mode, la_dtype val, la_dtype vec, const void *r, void *z,
                hipStream_t s);
void prec_apply_transpose(Prec &P, la_trsv_mode mode, la_dtype val, la_dtype vec,
                          const void *r, void *z, hipStream_t s);

// -------------------------------------------------------------- BLAS-1
// Deterministic FP64-accumulated reductions: a fixed grid of kRedBlocks
// blocks writes per-block partials, a one-block kernel sums them in order.
constexpr int kRedBlocks = 512;
struct Scratch {
  DevBuf partials;  // kRedBlocks * kMaxDots doubles
  DevBuf scalars;   // device scalars used by the solvers
  double *host = nullptr;  // pinned mirror of a few scalars
  // coherent pinned report slot: rep[0] holds the bits of the last sum, or
  // an empty marker the host writes before each report_sum; written by the
  // device and polled by the host (wait_report)
  uint64_t *rep = nullptr, *rep_dev = nullptr;
  uint64_t rep_seq = 0;
  uint64_t rep_pending = 0;  // ticket of the report not yet read, 0 if none
  Scratch();
  ~Scratch();
};
// the sum of kRedBlocks partials, stored as one word into the report slot
// (no copy command, no event); returns a ticket for wait_report. One report
// is outstanding at a time: a report_sum whose predecessor was never read
// waits for the stream first; wait_report rejects any ticket but the
// outstanding one.
uint64_t report_sum(Scratch &S, const double *part, hipStream_t s);
// waits (host spin, bounded) until the report has arrived; returns its value
double wait_report(Scratch &S, uint64_t seq, hipStream_t s);
constexpr int kMaxDots = 64;

// out[0] = <x,y>
void dot_dev(int64_t n, la_dtype vec, const void *x, const void *y, double *partials,
             double *out, hipStream_t s);
// out_dev[0] = <x,y> with the partial sums in a fixed per-device array: no
// allocation; calls that overlap on one GPU share that array and race
void dot_noalloc(int64_t n, la_dtype vec, const void *x, const void *y, double *out_dev,
                 hipStream_t s);
// out[i] = <V_i, w> for i < k; V is k vectors of length n, stride ld
void multidot_dev(int64_t n, int k, la_dtype vec, const void *V, int64_t ld, const void *w,
                  double *partials, double *out, hipStream_t s);
// w -= sum_i h[i] V_i (h device FP64)
void multiaxpy_dev(int64_t n, int k, la_dtype vec, const void *V, int64_t ld, const double *h,
                   void *w, hipStream_t s);
// y = a*x + b*y with host scalars
void axpby(int64_t n, la_dtype vec, double a, const void *x, double b, void *y,
           hipStream_t s);
void copy_vec(int64_t n, la_dtype src_dt, const void *src, la_dtype dst_dt, void *dst,
              double scale, hipStream_t s);
void fill_zero(int64_t n, la_dtype vec, void *x, hipStream_t s);

// CG step kernels. Dot products stay as kRedBlocks per-block partials; the
// consuming kernel sums them itself (same order in every block), and
// report_sum hands the host the same sum when it needs the value.
void dot_partials(int64_t n, la_dtype vec, const void *x, const void *y, double *partials,
                  hipStream_t s);
// x += a p, r -= a q, a = <r,z>/<p,q>; partials of <r,r> into part_rr
void cg_update_xr(int64_t n, la_dtype vec, void *x, void *r, const void *p, const void *q,
                  const double *part_rz, const double *part_pq, double *part_rr, hipStream_t s);
// p = z + (<r,z>_new / <r,z>_old) p
void cg_update_p(int64_t n, la_dtype vec, void *p, const void *z, const double *part_new,
                 const double *part_old, hipStream_t s);
// x64 += scale * d (d in vec precision)
void add_scaled(int64_t n, la_dtype dvec, const void *d, double scale, double *x64,
                hipStream_t s);
// v = w / h (h device scalar) ; used by GMRES
void scale_by_inv(int64_t n, la_dtype vec, const void *w, const double *h, void *v,
                  hipStream_t s);

}  // namespace laamd
