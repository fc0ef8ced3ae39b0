// extern "C" boundary of the library (declared in include/laamd.h). Every
// entry point catches all exceptions and returns an la_status; the message
// of the last failure is kept per thread for la_last_error().
#include <cstring>
#include <string>

#include "device.hpp"
#include "host_algos.hpp"

namespace laamd {
int pcg_solve(Mat &A, Prec *P, const la_solver_opts &o, const double *b, double *x,
              la_solver_result *res);
int fgmres_solve(Mat &A, Prec *P, const la_solver_opts &o, const double *b, double *x,
                 la_solver_result *res);
int pcg_solve_transpose(Mat &A, Prec *P, const la_solver_opts &o, const double *b, double *x,
                        la_solver_result *res);
int fgmres_solve_right(Mat &A, Prec *P, Prec *R, const la_solver_opts &o, const double *b,
                       double *x, la_solver_result *res);
int fgmres_solve_transpose(Mat &A, Prec *P, const la_solver_opts &o, const double *b, double *x,
                           la_solver_result *res);
void gen_uniform(int64_t n, uint64_t seed, double lo, double hi, double *out);
void gen_lognormal(int64_t n, uint64_t seed, double mu, double sigma, double *out);
void gen_poisson2d(int nx, int ny, int *ai, int *aj, double *av, int64_t *nnz);
void gen_laplace3d(int nx, int ny, int nz, int *ai, int *aj, double *av, int64_t *nnz);
void gen_powerlaw(int n, double mean_len, double sigma, int max_len, uint64_t seed, int *ai,
                  int *aj, double *av, int64_t *nnz);
void gen_aniso27(int nx, int ny, int nz, double aniso, uint64_t seed, int *ai, int *aj,
                 double *av, int64_t *nnz);
void gen_diffusion2d(int nx, int ny, const double *theta, int *ai, int *aj, double *av,
                     int *edge_i, int *edge_j, int *val_map, int64_t *nnz);

namespace {
thread_local std::string tl_err;
}
void set_last_error(const std::string &m) { tl_err = m; }
}  // namespace laamd

using namespace laamd;

struct la_mat_s {
  std::unique_ptr<Mat> m;
};
struct la_trsv_s {
  std::unique_ptr<TrsvPlan> t;
};
struct la_prec_s {
  std::unique_ptr<Prec> p;
};
struct la_csr_file_s {
  int rows = 0, cols = 0;
  std::vector<int> ai, aj;
  std::vector<double> av;
};
namespace laamd {
void read_matrix_market(const char *path, int base, int *rows, int *cols, std::vector<int> &ai,
                        std::vector<int> &aj, std::vector<double> &av);
void read_binary_coo_csr(const char *path, int base, int *rows, int *cols, std::vector<int> &ai,
                         std::vector<int> &aj, std::vector<double> &av);
void read_binary_vector(const char *path, std::vector<double> &v);
}  // namespace laamd

#define LA_TRY try {
#define LA_CATCH                                                   \
  }                                                                \
  catch (const la_error &e) {                                      \
    set_last_error(e.what());                                      \
    return e.code;                                                 \
  }                                                                \
  catch (const std::bad_alloc &) {                                 \
    set_last_error("host allocation failed");                      \
    return LA_ERR_ALLOC;                                           \
  }                                                                \
  catch (const std::exception &e) {                                \
    set_last_error(e.what());                                      \
    return LA_ERR_ARG;                                             \
  }                                                                \
  return LA_OK;

extern "C" {

const char *la_last_error(void) { return tl_err.c_str(); }
int la_version(void) { return 100; }

int la_device_alloc(size_t bytes, void **ptr) {
  LA_TRY
  LA_CHECK_ARG(ptr, "null output pointer");
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
  if (e != hipSuccess) throw la_error(LA_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
  LA_CATCH
}
int la_device_free(void *ptr) {
  LA_TRY
  if (ptr) LA_HIP(hipFree(ptr));
  LA_CATCH
}
int la_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  LA_TRY
  if (bytes) LA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  LA_CATCH
}
int la_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  LA_TRY
  if (bytes) LA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  LA_CATCH
}
int la_device_synchronize(void) {
  LA_TRY
  LA_HIP(hipDeviceSynchronize());
  LA_CATCH
}

// ------------------------------------------------------------ host algos
int la_topo_sort2(int n, int base, const int *ai, const int *aj, int upper, int *iperm,
                  int *prefix, int *nlevels) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && iperm && prefix, "null argument");
  const int L = topo_sort2(n, base, ai, aj, upper != 0, iperm, prefix);
  if (nlevels) *nlevels = L;
  LA_CATCH
}

int la_split_ldu(int n, int base, const int *ai, const int *aj, const double *av, int *Lai,
                 int *Laj, double *Lav, double *D, int *Uai, int *Uaj, double *Uav, int *lnnz,
                 int *unnz) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && Lai && Uai, "null argument");
  int64_t l = 0, u = 0;
  split_ldu(n, base, ai, aj, av, Lai, Laj, Lav, D, Uai, Uaj, Uav, &l, &u);
  if (lnnz) *lnnz = static_cast<int>(l);
  if (unnz) *unnz = static_cast<int>(u);
  LA_CATCH
}

int la_split_triangle(int n, int base, const int *ai, const int *aj, const double *av, int upper,
                      int *Tai, int *Taj, double *Tav, int *tnnz) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && Tai, "null argument");
  int64_t t = 0;
  split_triangle(n, base, ai, aj, av, upper != 0, Tai, Taj, Tav, &t);
  if (tnnz) *tnnz = static_cast<int>(t);
  LA_CATCH
}

int la_diagonal(int n, int base, const int *ai, const int *aj, const double *av, int *diagpos,
                double *diag, int *all_present) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai, "null argument");
  const bool ok = diagonal(n, base, ai, aj, av, diagpos, diag);
  if (all_present) *all_present = ok ? 1 : 0;
  LA_CATCH
}

int la_transpose(int rows, int cols, int base, const int *ai, const int *aj, const double *av,
                 int *ti, int *tj, double *tv) {
  LA_TRY
  LA_CHECK_ARG(rows >= 0 && cols >= 0 && ai && ti, "null argument");
  transpose(rows, cols, base, ai, aj, av, ti, tj, tv);
  LA_CATCH
}

int la_permute(int n, int base, const int *ai, const int *aj, const double *av, const int *iperm,
               const int *perm, int *pai, int *paj, double *pav) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && pai, "null argument");
  permute(n, base, ai, aj, av, iperm, perm, pai, paj, pav);
  LA_CATCH
}

int la_rcm(int n, int base, const int *ai, const int *aj, int *iperm, int *perm) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && iperm && perm, "null argument");
  rcm(n, base, ai, aj, iperm, perm);
  LA_CATCH
}
int la_rcm_threads(int n, int base, const int *ai, const int *aj, int nthreads, int *iperm,
                   int *perm) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && iperm && perm, "null argument");
  rcm(n, base, ai, aj, iperm, perm, nthreads);
  LA_CATCH
}

int la_iluk_symbolic(int n, int base, const int *ai, const int *aj, int level, int *fai,
                     int *faj, int *first_upper, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && fai && level >= 0, "bad argument");
  std::vector<int> vi, vj, fu;
  iluk_symbolic(n, base, ai, aj, level, vi, vj, fu);
  std::memcpy(fai, vi.data(), sizeof(int) * (n + 1));
  if (nnz) *nnz = static_cast<int64_t>(vj.size());
  if (faj) std::memcpy(faj, vj.data(), sizeof(int) * vj.size());
  if (first_upper) std::memcpy(first_upper, fu.data(), sizeof(int) * n);
  LA_CATCH
}

int la_iluk_numeric(int n, int base, const int *ai, const int *aj, const double *av,
                    const int *fai, const int *faj, const int *first_upper, double *fav) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && fai && first_upper && fav, "null argument");
  if (!iluk_numeric(n, base, ai, aj, av, fai, faj, first_upper, fav))
    throw la_error(LA_ERR_NUMERIC, "no element on diagonal");
  LA_CATCH
}

int la_icc_symbolic(int n, int base, const int *ai, const int *aj, const int *diag_pos, int lvl,
                    int *icc_ai, int *icc_aj, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && diag_pos && icc_ai && lvl >= 0, "bad argument");
  std::vector<int> vi, vj;
  icc_symbolic(n, base, ai, aj, diag_pos, lvl, vi, vj);
  std::memcpy(icc_ai, vi.data(), sizeof(int) * (n + 1));
  if (nnz) *nnz = static_cast<int64_t>(vj.size());
  if (icc_aj) std::memcpy(icc_aj, vj.data(), sizeof(int) * vj.size());
  LA_CATCH
}

int la_icc_numeric(int n, int base, const int *ai, const int *aj, const double *av,
                   const int *diag_pos, double omega, const int *icc_ai, const int *icc_aj,
                   double *icc_av, int *ok) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && diag_pos && icc_ai && icc_av, "null argument");
  LA_CHECK_ARG(icc_ai[n] == icc_ai[0] || (aj && av && icc_aj), "null argument");
  const bool good = icc_numeric(n, base, ai, aj, av, diag_pos, omega, icc_ai, icc_aj, icc_av);
  if (ok) *ok = good ? 1 : 0;
  LA_CATCH
}

int la_ick_numeric(int n, int base, const int *ai, const int *aj, const double *av,
                   const int *diag_pos, double shift, const int *icc_ai, const int *icc_aj,
                   double *icc_av, int *ok) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && diag_pos && icc_ai && icc_av, "null argument");
  LA_CHECK_ARG(icc_ai[n] == icc_ai[0] || (aj && av && icc_aj), "null argument");
  const bool good = ick_numeric(n, base, ai, aj, av, diag_pos, shift, icc_ai, icc_aj, icc_av);
  if (ok) *ok = good ? 1 : 0;
  LA_CATCH
}

// ------------------------------------------------------------ generators
int la_gen_poisson2d(int nx, int ny, int *ai, int *aj, double *av, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(nx > 0 && ny > 0 && ai, "bad argument");
  gen_poisson2d(nx, ny, ai, aj, av, nnz);
  LA_CATCH
}
int la_gen_laplace3d(int nx, int ny, int nz, int *ai, int *aj, double *av, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(nx > 0 && ny > 0 && nz > 0 && ai, "bad argument");
  gen_laplace3d(nx, ny, nz, ai, aj, av, nnz);
  LA_CATCH
}
int la_gen_powerlaw(int n, double mean_len, double sigma, int max_len, uint64_t seed, int *ai,
                    int *aj, double *av, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(n > 0 && mean_len >= 1 && max_len >= 1 && ai, "bad argument");
  gen_powerlaw(n, mean_len, sigma, max_len, seed, ai, aj, av, nnz);
  LA_CATCH
}
int la_gen_aniso27(int nx, int ny, int nz, double aniso, uint64_t seed, int *ai, int *aj,
                   double *av, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(nx > 0 && ny > 0 && nz > 0 && aniso > 0 && ai, "bad argument");
  gen_aniso27(nx, ny, nz, aniso, seed, ai, aj, av, nnz);
  LA_CATCH
}
int la_gen_diffusion2d(int nx, int ny, const double *theta, int *ai, int *aj, double *av,
                       int *edge_i, int *edge_j, int *val_map, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(nx > 1 && ny > 1 && theta && ai, "bad argument");
  LA_CHECK_ARG(!aj || (av && edge_i && edge_j && val_map), "null output");
  gen_diffusion2d(nx, ny, theta, ai, aj, av, edge_i, edge_j, val_map, nnz);
  LA_CATCH
}
int la_gen_uniform(int64_t n, uint64_t seed, double lo, double hi, double *out) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && (n == 0 || out), "bad argument");
  gen_uniform(n, seed, lo, hi, out);
  LA_CATCH
}
int la_gen_lognormal(int64_t n, uint64_t seed, double mu, double sigma, double *out) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && (n == 0 || out), "bad argument");
  gen_lognormal(n, seed, mu, sigma, out);
  LA_CATCH
}

// ------------------------------------------------------------ matrix
int la_mat_create(int rows, int cols, int base, const int *ai, const int *aj, const double *av,
                  la_mat_t *out) {
  LA_TRY
  LA_CHECK_ARG(out, "null output handle");
  LA_CHECK_ARG(base == 0 || base == 1, "base must be 0 or 1");
  auto h = std::make_unique<la_mat_s>();
  h->m = make_mat(rows, cols, base, ai, aj, av, nullptr);
  *out = h.release();
  LA_CATCH
}
int la_mat_destroy(la_mat_t A) {
  LA_TRY
  delete A;
  LA_CATCH
}
int la_mat_add_dtype(la_mat_t A, la_dtype dt) {
  LA_TRY
  LA_CHECK_ARG(A && dt >= LA_F64 && dt <= LA_BF16, "bad argument");
  A->m->vals.add(dt, A->m->stream);
  LA_HIP(hipStreamSynchronize(A->m->stream));
  LA_CATCH
}
int la_mat_set_values(la_mat_t A, const double *av) {
  LA_TRY
  LA_CHECK_ARG(A && (av || A->m->nnz == 0), "bad argument");
  A->m->vals.set_from_host(av, A->m->nnz, A->m->stream);
  LA_HIP(hipStreamSynchronize(A->m->stream));
  LA_CATCH
}
int la_mat_set_values_device(la_mat_t A, const double *av) {
  LA_TRY
  LA_CHECK_ARG(A && (av || A->m->nnz == 0), "bad argument");
  A->m->vals.set_from_device_f64(av, A->m->nnz, A->m->stream);
  LA_CATCH
}
int la_mat_set_stream(la_mat_t A, void *s) {
  LA_TRY
  LA_CHECK_ARG(A, "null handle");
  A->m->stream = static_cast<hipStream_t>(s);
  if (A->m->transposed) A->m->transposed->stream = static_cast<hipStream_t>(s);
  LA_CATCH
}
int la_mat_info(la_mat_t A, int *rows, int *cols, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(A, "null handle");
  if (rows) *rows = A->m->rows;
  if (cols) *cols = A->m->cols;
  if (nnz) *nnz = A->m->nnz;
  LA_CATCH
}
int la_spmv_preprocess(la_mat_t A, la_spmv_alg alg) {
  LA_TRY
  LA_CHECK_ARG(A, "null handle");
  LA_CHECK_ARG(alg >= LA_SPMV_ROW && alg <= LA_SPMV_COLBLOCK, "bad SpMV algorithm");
  if (alg == LA_SPMV_SEGSUM || alg == LA_SPMV_ALBUS || alg == LA_SPMV_AUTO)
    spmv_prepare_segments(*A->m);
  if (alg == LA_SPMV_SELL || alg == LA_SPMV_AUTO || alg == LA_SPMV_COLBLOCK)
    (void)spmv_resolve(*A->m, alg);
  LA_CATCH
}
int la_spmv_resolve(la_mat_t A, la_spmv_alg alg, int *chosen) {
  LA_TRY
  LA_CHECK_ARG(A && chosen, "bad argument");
  LA_CHECK_ARG(alg >= LA_SPMV_ROW && alg <= LA_SPMV_COLBLOCK, "bad SpMV algorithm");
  *chosen = static_cast<int>(spmv_resolve(*A->m, alg));
  LA_CATCH
}
int la_mat_sell_info(la_mat_t A, int *built, int *col16, int64_t *slots) {
  LA_TRY
  LA_CHECK_ARG(A, "null handle");
  const bool ok = spmv_prepare_sell(*A->m);
  if (built) *built = ok ? 1 : 0;
  if (col16) *col16 = ok && A->m->sell_col16 ? 1 : 0;
  if (slots) *slots = ok ? A->m->sell_len : 0;
  LA_CATCH
}
int la_spmv(la_mat_t A, la_spmv_alg alg, la_dtype val, la_dtype vec, la_dtype acc, double alpha,
            const void *x, double beta, void *y) {
  LA_TRY
  LA_CHECK_ARG(A, "null handle");
  spmv(*A->m, alg, val, vec, acc, alpha, x, beta, y, y, A->m->stream);
  LA_CATCH
}
int la_spmv_axpby(la_mat_t A, la_spmv_alg alg, la_dtype val, la_dtype vec, la_dtype acc,
                  double alpha, const void *x, double beta, const void *z, void *y) {
  LA_TRY
  LA_CHECK_ARG(A && (beta == 0.0 || z), "bad argument");
  spmv(*A->m, alg, val, vec, acc, alpha, x, beta, z, y, A->m->stream);
  LA_CATCH
}
int la_spmv_transpose(la_mat_t A, la_dtype val, la_dtype vec, double alpha, const void *x,
                      double beta, void *y) {
  LA_TRY
  LA_CHECK_ARG(A, "null handle");
  spmv_transpose(*A->m, val, vec, alpha, x, beta, y, A->m->stream);
  LA_CATCH
}
int la_csr_value_grad(la_mat_t A, const double *lambda, const double *x, double *g) {
  LA_TRY
  LA_CHECK_ARG(A && lambda && x && g, "null argument");
  csr_value_grad(*A->m, lambda, x, g, A->m->stream);
  LA_CATCH
}

// ------------------------------------------------------------ trsv
int la_trsv_create(int n, int base, const int *ai, const int *aj, const double *av,
                   const double *diag, la_uplo uplo, la_trsv_t *out) {
  LA_TRY
  LA_CHECK_ARG(out && (base == 0 || base == 1), "bad argument");
  LA_CHECK_ARG(uplo == LA_LOWER || uplo == LA_UPPER, "bad uplo");
  auto h = std::make_unique<la_trsv_s>();
  h->t = make_trsv(n, base, ai, aj, av, diag, uplo == LA_UPPER, nullptr);
  *out = h.release();
  LA_CATCH
}
int la_trsv_destroy(la_trsv_t T) {
  LA_TRY
  delete T;
  LA_CATCH
}
int la_trsv_add_dtype(la_trsv_t T, la_dtype dt) {
  LA_TRY
  LA_CHECK_ARG(T && dt >= LA_F64 && dt <= LA_BF16, "bad argument");
  trsv_add_dtype(*T->t, dt);
  LA_CATCH
}
int la_trsv_set_stream(la_trsv_t T, void *s) {
  LA_TRY
  LA_CHECK_ARG(T, "null handle");
  T->t->stream = static_cast<hipStream_t>(s);
  if (T->t->transposed) T->t->transposed->stream = static_cast<hipStream_t>(s);
  LA_CATCH
}
int la_trsv_info(la_trsv_t T, int *n, int *nlevels, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(T, "null handle");
  if (n) *n = T->t->n;
  if (nlevels) *nlevels = T->t->nlevels;
  if (nnz) *nnz = T->t->nnz;
  LA_CATCH
}
int la_trsv_set_sweeps(la_trsv_t T, int sweeps) {
  LA_TRY
  LA_CHECK_ARG(T, "null handle");
  trsv_set_sweeps(*T->t, sweeps);
  LA_CATCH
}
int la_trsv_solve(la_trsv_t T, la_trsv_mode mode, la_dtype val, la_dtype vec, const void *b,
                  void *x) {
  LA_TRY
  LA_CHECK_ARG(T, "null handle");
  trsv_solve(*T->t, mode, val, vec, b, x, T->t->stream);
  if (mode == LA_TRSV_SYNCFREE && trsv_check_timeout(*T->t, T->t->stream))
    throw la_error(LA_ERR_TIMEOUT, "sync-free triangular solve exceeded its spin bound");
  LA_CATCH
}
int la_trsv_solve_transpose(la_trsv_t T, la_trsv_mode mode, la_dtype val, la_dtype vec,
                            const void *b, void *x) {
  LA_TRY
  LA_CHECK_ARG(T, "null handle");
  TrsvPlan &Tt = trsv_transposed(*T->t);
  trsv_solve(Tt, mode, val, vec, b, x, T->t->stream);
  if (mode == LA_TRSV_SYNCFREE && trsv_check_timeout(Tt, T->t->stream))
    throw la_error(LA_ERR_TIMEOUT, "sync-free triangular solve exceeded its spin bound");
  LA_CATCH
}

// ------------------------------------------------------------ preconditioner
int la_prec_create_ilu(int n, int base, const int *fai, const int *faj, const double *fav,
                       la_prec_t *out) {
  LA_TRY
  LA_CHECK_ARG(out && fai && (base == 0 || base == 1), "bad argument");
  auto h = std::make_unique<la_prec_s>();
  h->p = make_prec_ilu(n, base, fai, faj, fav, nullptr);
  *out = h.release();
  LA_CATCH
}
int la_prec_create_ic(int n, int base, const int *rai, const int *raj, const double *rav,
                      la_prec_t *out) {
  LA_TRY
  LA_CHECK_ARG(out && rai && (base == 0 || base == 1), "bad argument");
  auto h = std::make_unique<la_prec_s>();
  h->p = make_prec_ic(n, base, rai, raj, rav, nullptr);
  *out = h.release();
  LA_CATCH
}
int la_prec_create_ilu_gpu(int n, int base, const int *ai, const int *aj, int level,
                           la_prec_t *out) {
  LA_TRY
  LA_CHECK_ARG(out && ai && (base == 0 || base == 1) && n >= 0 && level >= 0, "bad argument");
  LA_CHECK_ARG(ai[0] == base && (ai[n] == base || aj), "bad CSR");
  auto h = std::make_unique<la_prec_s>();
  h->p = make_prec_ilu_gpu(n, base, ai, aj, level, nullptr);
  *out = h.release();
  LA_CATCH
}
int la_prec_create_ic_gpu(int n, int base, const int *ai, const int *aj, int level, double omega,
                          la_prec_t *out) {
  LA_TRY
  LA_CHECK_ARG(out && ai && (base == 0 || base == 1) && n >= 0 && level >= 0, "bad argument");
  LA_CHECK_ARG(ai[0] == base && (ai[n] == base || aj), "bad CSR");
  auto h = std::make_unique<la_prec_s>();
  h->p = make_prec_ic_gpu(n, base, ai, aj, level, omega, nullptr);
  *out = h.release();
  LA_CATCH
}
int la_prec_create_ick_gpu(int n, int base, const int *ai, const int *aj, int level, double shift,
                           la_prec_t *out) {
  LA_TRY
  LA_CHECK_ARG(out && ai && (base == 0 || base == 1) && n >= 0 && level >= 0, "bad argument");
  LA_CHECK_ARG(ai[0] == base && (ai[n] == base || aj), "bad CSR");
  auto h = std::make_unique<la_prec_s>();
  h->p = make_prec_ic_gpu(n, base, ai, aj, level, shift, nullptr, true);
  *out = h.release();
  LA_CATCH
}
int la_prec_refactor(la_prec_t P, const double *a_vals) {
  LA_TRY
  LA_CHECK_ARG(P, "null handle");
  prec_refactor(*P->p, a_vals, P->p->stream);
  LA_CATCH
}
int la_prec_factor(la_prec_t P, int *fai, int *faj, double *fav, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(P && nnz, "bad argument");
  std::vector<int> i, j;
  std::vector<double> v;
  prec_factor(*P->p, &i, &j, fav ? &v : nullptr, P->p->stream);
  *nnz = static_cast<int64_t>(j.size());
  if (fai) std::copy(i.begin(), i.end(), fai);
  if (faj) std::copy(j.begin(), j.end(), faj);
  if (fav) std::copy(v.begin(), v.end(), fav);
  LA_CATCH
}
int la_prec_destroy(la_prec_t P) {
  LA_TRY
  delete P;
  LA_CATCH
}
int la_prec_add_dtype(la_prec_t P, la_dtype dt) {
  LA_TRY
  LA_CHECK_ARG(P && dt >= LA_F64 && dt <= LA_BF16, "bad argument");
  prec_add_dtype(*P->p, dt);
  LA_CATCH
}
int la_prec_set_stream(la_prec_t P, void *s) {
  LA_TRY
  LA_CHECK_ARG(P, "null handle");
  hipStream_t st = static_cast<hipStream_t>(s);
  P->p->stream = st;
  P->p->lower->stream = st;
  P->p->upper->stream = st;
  LA_CATCH
}
int la_prec_apply(la_prec_t P, la_trsv_mode mode, la_dtype val, la_dtype vec, const void *r,
                  void *z) {
  LA_TRY
  LA_CHECK_ARG(P, "null handle");
  prec_apply(*P->p, mode, val, vec, r, z, P->p->stream);
  if (mode == LA_TRSV_SYNCFREE && (trsv_check_timeout(*P->p->lower, P->p->stream) ||
                                   trsv_check_timeout(*P->p->upper, P->p->stream)))
    throw la_error(LA_ERR_TIMEOUT, "sync-free triangular solve exceeded its spin bound");
  LA_CATCH
}
int la_prec_apply_transpose(la_prec_t P, la_trsv_mode mode, la_dtype val, la_dtype vec,
                            const void *r, void *z) {
  LA_TRY
  LA_CHECK_ARG(P, "null handle");
  prec_apply_transpose(*P->p, mode, val, vec, r, z, P->p->stream);
  LA_CATCH
}
int la_prec_info(la_prec_t P, int *levels_lower, int *levels_upper) {
  LA_TRY
  LA_CHECK_ARG(P, "null handle");
  if (levels_lower) *levels_lower = P->p->lower->nlevels;
  if (levels_upper) *levels_upper = P->p->upper->nlevels;
  LA_CATCH
}
int la_trsv_schedule_info(la_trsv_t T, int *tiled, int *tiles, int *tile_edge) {
  LA_TRY
  LA_CHECK_ARG(T, "null handle");
  if (tiled) *tiled = T->t->tiled ? 1 : 0;
  if (tiles) *tiles = T->t->t_nparts;
  if (tile_edge) *tile_edge = T->t->t_tau;
  LA_CATCH
}
int la_prec_schedule_info(la_prec_t P, int *tiled_lower, int *tiled_upper) {
  LA_TRY
  LA_CHECK_ARG(P, "null handle");
  if (tiled_lower) *tiled_lower = P->p->lower->tiled ? P->p->lower->t_nparts : 0;
  if (tiled_upper) *tiled_upper = P->p->upper->tiled ? P->p->upper->t_nparts : 0;
  LA_CATCH
}
int la_prec_set_sweeps(la_prec_t P, int sweeps) {
  LA_TRY
  LA_CHECK_ARG(P, "null handle");
  trsv_set_sweeps(*P->p->lower, sweeps);
  trsv_set_sweeps(*P->p->upper, sweeps);
  LA_CATCH
}

// ------------------------------------------------------------ BLAS-1
namespace {
Scratch &blas_scratch() {
  static thread_local std::unique_ptr<Scratch> s;
  if (!s) s = std::make_unique<Scratch>();
  return *s;
}
}  // namespace

int la_dot(int64_t n, la_dtype vec, const void *x, const void *y, double *result,
           void *stream) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && (vec == LA_F64 || vec == LA_F32) && result, "bad argument");
  LA_CHECK_ARG(n == 0 || (x && y), "null vector");
  hipStream_t s = static_cast<hipStream_t>(stream);
  Scratch &S = blas_scratch();
  dot_dev(n, vec, x, y, S.partials.as<double>(), S.scalars.as<double>(), s);
  LA_HIP(hipMemcpyAsync(S.host, S.scalars.p, sizeof(double), hipMemcpyDeviceToHost, s));
  LA_HIP(hipStreamSynchronize(s));
  *result = S.host[0];
  LA_CATCH
}

int la_dot_noalloc(int64_t n, la_dtype vec, const void *x, const void *y, double *result_dev,
                   void *stream) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && (vec == LA_F64 || vec == LA_F32) && result_dev, "bad argument");
  LA_CHECK_ARG(n == 0 || (x && y), "null vector");
  dot_noalloc(n, vec, x, y, result_dev, static_cast<hipStream_t>(stream));
  LA_CATCH
}

int la_axpy_dot(int64_t n, la_dtype vec, double a, const void *x, void *y, const void *z,
                double *result, void *stream) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && (vec == LA_F64 || vec == LA_F32) && result, "bad argument");
  LA_CHECK_ARG(n == 0 || (x && y), "null vector");
  hipStream_t s = static_cast<hipStream_t>(stream);
  axpby(n, vec, a, x, 1.0, y, s);
  Scratch &S = blas_scratch();
  dot_dev(n, vec, y, z ? z : y, S.partials.as<double>(), S.scalars.as<double>(), s);
  LA_HIP(hipMemcpyAsync(S.host, S.scalars.p, sizeof(double), hipMemcpyDeviceToHost, s));
  LA_HIP(hipStreamSynchronize(s));
  *result = S.host[0];
  LA_CATCH
}

// ------------------------------------------------------------ solvers
void la_solver_opts_default(la_solver_opts *o) {
  if (!o) return;
  o->rel_tol = 1e-8;
  o->abs_tol = 0.0;
  o->max_iters = 1000;
  o->restart = 30;
  o->stop_rule = LA_STOP_TRUE_RESIDUAL;
  o->precision = LA_PREC_FP64;
  o->trsv_mode = LA_TRSV_SYNCFREE;
  o->spmv_alg = LA_SPMV_AUTO;
  o->inner_tol = 1e-5;
  o->max_outer = 20;
  o->print_level = 0;
  o->reliable_delta = 0.1;
}

int la_pcg_solve(la_mat_t A, la_prec_t P, const la_solver_opts *opts, const double *b,
                 double *x, la_solver_result *res) {
  LA_TRY
  LA_CHECK_ARG(A && opts && b && x, "null argument");
  if (P) P->p->lower->stream = P->p->upper->stream = A->m->stream;
  const int st = pcg_solve(*A->m, P ? P->p.get() : nullptr, *opts, b, x, res);
  if (st != LA_OK) {
    set_last_error("PCG did not reach the tolerance within the iteration limit");
    return st;
  }
  LA_CATCH
}

int la_fgmres_solve(la_mat_t A, la_prec_t P, const la_solver_opts *opts, const double *b,
                    double *x, la_solver_result *res) {
  LA_TRY
  LA_CHECK_ARG(A && opts && b && x, "null argument");
  if (P) P->p->lower->stream = P->p->upper->stream = A->m->stream;
  const int st = fgmres_solve(*A->m, P ? P->p.get() : nullptr, *opts, b, x, res);
  if (st != LA_OK) {
    set_last_error("FGMRES did not reach the tolerance within the iteration limit");
    return st;
  }
  LA_CATCH
}

int la_pcg_solve_transpose(la_mat_t A, la_prec_t P, const la_solver_opts *opts, const double *b,
                           double *x, la_solver_result *res) {
  LA_TRY
  LA_CHECK_ARG(A && opts && b && x, "null argument");
  if (P) P->p->lower->stream = P->p->upper->stream = A->m->stream;
  const int st = pcg_solve_transpose(*A->m, P ? P->p.get() : nullptr, *opts, b, x, res);
  if (st != LA_OK) {
    set_last_error("PCG (transposed) did not reach the tolerance within the iteration limit");
    return st;
  }
  LA_CATCH
}

int la_fgmres_solve_right(la_mat_t A, la_prec_t P, la_prec_t R, const la_solver_opts *opts,
                          const double *b, double *x, la_solver_result *res) {
  LA_TRY
  LA_CHECK_ARG(A && R && opts && b && x, "bad argument");
  const int st = fgmres_solve_right(*A->m, P ? P->p.get() : nullptr, R->p.get(), *opts, b, x, res);
  if (st == LA_ERR_NOT_CONVERGED) {
    set_last_error("FGMRES (right operator) did not reach the tolerance within the iteration limit");
    return st;
  }
  LA_CATCH
}
int la_fgmres_solve_transpose(la_mat_t A, la_prec_t P, const la_solver_opts *opts,
                              const double *b, double *x, la_solver_result *res) {
  LA_TRY
  LA_CHECK_ARG(A && opts && b && x, "null argument");
  if (P) P->p->lower->stream = P->p->upper->stream = A->m->stream;
  const int st = fgmres_solve_transpose(*A->m, P ? P->p.get() : nullptr, *opts, b, x, res);
  if (st != LA_OK) {
    set_last_error("FGMRES (transposed) did not reach the tolerance within the iteration limit");
    return st;
  }
  LA_CATCH
}

// ------------------------------------------------------------ scaling
int la_rowwise_sqrt_norm(int n, int base, const int *ai, const double *av, double *d) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && d && (base == 0 || base == 1), "bad argument");
  LA_CHECK_ARG(ai[n] == ai[0] || av, "null values");
  rowwise_sqrt_norm(n, base, ai, av, d);
  LA_CATCH
}
int la_dtad(int n, int base, const int *ai, const int *aj, double *av, const double *d) {
  LA_TRY
  LA_CHECK_ARG(n >= 0 && ai && d && (base == 0 || base == 1), "bad argument");
  LA_CHECK_ARG(ai[n] == ai[0] || (aj && av), "null arrays");
  for (int64_t q = 0; q < static_cast<int64_t>(ai[n]) - ai[0]; ++q)
    LA_CHECK_ARG(aj[q] - base >= 0 && aj[q] - base < n, "column out of range");
  dtad(n, base, ai, aj, av, d);
  LA_CATCH
}
int la_mat_rowwise_sqrt_norm(la_mat_t A, double *d) {
  LA_TRY
  LA_CHECK_ARG(A && (d || A->m->rows == 0), "bad argument");
  mat_rowwise_sqrt_norm(*A->m, d, A->m->stream);
  LA_HIP(hipStreamSynchronize(A->m->stream));
  LA_CATCH
}
int la_mat_dtad(la_mat_t A, const double *d) {
  LA_TRY
  LA_CHECK_ARG(A && (d || A->m->rows == 0), "bad argument");
  mat_dtad(*A->m, d, A->m->stream);
  LA_HIP(hipStreamSynchronize(A->m->stream));
  LA_CATCH
}

// ------------------------------------------------------------ readers
int la_read_matrix_market(const char *path, int base, la_csr_file_t *out) {
  LA_TRY
  LA_CHECK_ARG(out, "null output");
  auto h = std::make_unique<la_csr_file_s>();
  read_matrix_market(path, base, &h->rows, &h->cols, h->ai, h->aj, h->av);
  *out = h.release();
  LA_CATCH
}
int la_read_binary_coo(const char *path, int base, la_csr_file_t *out) {
  LA_TRY
  LA_CHECK_ARG(out, "null output");
  auto h = std::make_unique<la_csr_file_s>();
  read_binary_coo_csr(path, base, &h->rows, &h->cols, h->ai, h->aj, h->av);
  *out = h.release();
  LA_CATCH
}
int la_csr_file_info(la_csr_file_t F, int *rows, int *cols, int64_t *nnz) {
  LA_TRY
  LA_CHECK_ARG(F, "null handle");
  if (rows) *rows = F->rows;
  if (cols) *cols = F->cols;
  if (nnz) *nnz = static_cast<int64_t>(F->aj.size());
  LA_CATCH
}
int la_csr_file_copy(la_csr_file_t F, int *ai, int *aj, double *av) {
  LA_TRY
  LA_CHECK_ARG(F, "null handle");
  if (ai) std::copy(F->ai.begin(), F->ai.end(), ai);
  if (aj) std::copy(F->aj.begin(), F->aj.end(), aj);
  if (av) std::copy(F->av.begin(), F->av.end(), av);
  LA_CATCH
}
int la_csr_file_free(la_csr_file_t F) {
  LA_TRY
  delete F;
  LA_CATCH
}
int la_read_binary_vector(const char *path, double *out, int64_t *n) {
  LA_TRY
  LA_CHECK_ARG(n, "null size");
  std::vector<double> v;
  read_binary_vector(path, v);
  if (out) {
    LA_CHECK_ARG(*n >= static_cast<int64_t>(v.size()), "output too small");
    std::copy(v.begin(), v.end(), out);
  }
  *n = static_cast<int64_t>(v.size());
  LA_CATCH
}

}  // extern "C"
