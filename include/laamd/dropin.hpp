// C++ drop-in layer over the laamd C ABI (include/laamd.h).
//
// It keeps the reference's C++ call signatures so that callers of
// miaodi/linearAlgebra's hot path can switch by changing the include and
// the namespace (laamd::matrix_utils / laamd::mkl_wrapper in place of
// matrix_utils / mkl_wrapper):
//   SpMV functors and SPMV<CSR,Impl>          sparse_mat_op/spmv.hpp:12-334
//   triangular solves, OptimizedTriangularSolve sparse_mat_op/triangle_solve.hpp
//   TopologicalSort2 / SplitLDU / SplitTriangle / Diagonal / transpose /
//   permute                                    sparse_mat_op/matrix_utils.hpp
//   ICCLevelSymbolic0..3 / ICCLevelNumeric     sparse_mat_op/precond.hpp
//   incomplete_lu_k, incomplete_cholesky_k(_2) mkl_wrapper/incomplete_*.h
//   mkl_pcg_solver, mkl_fgmres_solver, solver_factory  mkl_wrapper/mkl_solver.h
// The arithmetic runs on the GPU through the C ABI; nothing here uses MKL.
// Host pointers in, host pointers out, as in the reference; vectors cross
// PCIe on every call of the host-pointer overloads, so loops that must stay
// on the device use the *_device overloads (device pointers).
//
// Errors: the reference reports failures as `false` plus text on stderr;
// this layer keeps the `bool` returns and throws laamd::error only where the
// reference would have had undefined behaviour (bad input) or the device
// failed.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../laamd.h"

namespace laamd {

struct error : std::runtime_error {
  int code;
  error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

namespace detail {
inline void check(int st) {
  if (st != LA_OK) throw error(st, std::string("laamd: ") + la_last_error());
}

// owning device buffer
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes) { resize(bytes); }
  ~DeviceBuffer() { if (p_) la_device_free(p_); }
  DeviceBuffer(const DeviceBuffer &) = delete;
  DeviceBuffer &operator=(const DeviceBuffer &) = delete;
  void resize(size_t bytes) {
    if (bytes <= bytes_ && p_) return;
    if (p_) la_device_free(p_);
    p_ = nullptr;
    check(la_device_alloc(bytes, &p_));
    bytes_ = bytes;
  }
  void *get() const { return p_; }
  void upload(const void *h, size_t b) { resize(b); check(la_memcpy_h2d(p_, h, b)); }
  void download(void *h, size_t b) const { check(la_memcpy_d2h(h, p_, b)); }

 private:
  void *p_ = nullptr;
  size_t bytes_ = 0;
};

struct MatHandle {
  la_mat_t h = nullptr;
  MatHandle(int rows, int cols, int base, const int *ai, const int *aj, const double *av) {
    check(la_mat_create(rows, cols, base, ai, aj, av, &h));
  }
  ~MatHandle() { if (h) la_mat_destroy(h); }
  MatHandle(const MatHandle &) = delete;
  MatHandle &operator=(const MatHandle &) = delete;
};

struct TrsvHandle {
  la_trsv_t h = nullptr;
  TrsvHandle(int n, int base, const int *ai, const int *aj, const double *av,
             const double *diag, la_uplo uplo) {
    check(la_trsv_create(n, base, ai, aj, av, diag, uplo, &h));
  }
  ~TrsvHandle() { if (h) la_trsv_destroy(h); }
  TrsvHandle(const TrsvHandle &) = delete;
  TrsvHandle &operator=(const TrsvHandle &) = delete;
};

// y = alpha*A*x + beta*y with host vectors, A given by host arrays
inline void host_spmv(la_mat_t A, la_spmv_alg alg, int rows, int cols, const double *b,
                      double *x, double alpha, double beta) {
  DeviceBuffer db(sizeof(double) * (cols ? cols : 1)), dx(sizeof(double) * (rows ? rows : 1));
  db.upload(b, sizeof(double) * cols);
  if (beta != 0.0) dx.upload(x, sizeof(double) * rows);
  check(la_spmv(A, alg, LA_F64, LA_F64, LA_F64, alpha, db.get(), beta, dx.get()));
  dx.download(x, sizeof(double) * rows);
}
}  // namespace detail

// =====================================================================
namespace matrix_utils {

enum TriangularMatrix { L = 0, U = 1 };
enum class FBSubstitutionType { Barrier, NoBarrier, NoBarrierSuperNode };

// sparse_mat_op/matrix_utils.hpp:83-103
template <typename R, typename C, typename V> struct CSRMatrixVec {
  using ROWTYPE = R;
  using COLTYPE = C;
  using VALTYPE = V;
  COLTYPE rows = 0;
  COLTYPE cols = 0;
  std::vector<ROWTYPE> ai;
  std::vector<COLTYPE> aj;
  std::vector<VALTYPE> av;
  ROWTYPE Base() const { return ai[0]; }
  ROWTYPE NNZ() const { return ai[rows] - ai[0]; }
  ROWTYPE const *AI() const { return ai.data(); }
  COLTYPE const *AJ() const { return aj.data(); }
  VALTYPE const *AV() const { return av.data(); }
};

// ------------------------------------------------------------- SpMV
// SerialSPMV / ParallelSPMV (spmv.hpp:12-57): x = alpha*A*b + beta*x. The
// free-form call re-uploads A; prefer SPMV<> or the preprocessing functors.
struct SerialSPMV {
  void operator()(int size, int base, int const *ai, int const *aj, double const *av,
                  double const *b, double *x, double alpha, double beta) const {
    detail::MatHandle A(size, size, base, ai, aj, av);
    detail::host_spmv(A.h, LA_SPMV_ROW, size, size, b, x, alpha, beta);
  }
};
struct ParallelSPMV : SerialSPMV {};

// SegSumSPMV / ALBUSSPMV (spmv.hpp:59-300): preprocess uploads the pattern
// and its nnz-balanced segment table once; operator() re-reads av on every
// call, as the reference does (spmv.hpp:230-293), so values updated in place
// between calls are used. The thread count of the reference has no meaning
// here (segments are fixed-size).
template <typename ROWTYPE = int, typename COLTYPE = int, typename VALTYPE = double,
          la_spmv_alg ALG = LA_SPMV_ALBUS>
class BalancedSPMV {
 public:
  explicit BalancedSPMV(int num_threads = 1) { (void)num_threads; }
  void setNumThreads(int) {}
  void preprocess(COLTYPE size, int base, ROWTYPE const *ai, COLTYPE const *aj,
                  VALTYPE const *av) {
    A_ = std::make_shared<detail::MatHandle>(size, size, base, ai, aj, av);
    detail::check(la_spmv_preprocess(A_->h, ALG));
    key_ = {ai, aj};
    size_ = size;
  }
  void operator()(COLTYPE size, int base, ROWTYPE const *ai, COLTYPE const *aj,
                  VALTYPE const *av, VALTYPE const *b, VALTYPE *x, VALTYPE alpha,
                  VALTYPE beta) const {
    if (!A_ || key_ != Key{ai, aj} || size != size_) {
      detail::MatHandle A(size, size, base, ai, aj, av);
      detail::host_spmv(A.h, ALG, size, size, b, x, alpha, beta);
      return;
    }
    detail::check(la_mat_set_values(A_->h, av));
    detail::host_spmv(A_->h, ALG, size, size, b, x, alpha, beta);
  }

 private:
  struct Key {
    const void *a, *b;
    bool operator!=(const Key &o) const { return a != o.a || b != o.b; }
  };
  std::shared_ptr<detail::MatHandle> A_;
  Key key_{nullptr, nullptr};
  COLTYPE size_ = 0;
};
template <typename R = int, typename C = int, typename V = double>
using SegSumSPMV = BalancedSPMV<R, C, V, LA_SPMV_SEGSUM>;
template <typename R = int, typename C = int, typename V = double>
using ALBUSSPMV = BalancedSPMV<R, C, V, LA_SPMV_ALBUS>;

// SPMV<CSR, Impl> (spmv.hpp:310-334). Like the reference, every product
// reads the matrix's current values (`_matrix->AV()`): the pattern is
// uploaded once, the values on each call. A caller whose values stay fixed
// can say so with values_constant(true) to skip that upload in
// apply_device (the device-resident form); the host form always uploads.
template <typename CSRMatrixType, typename SPMVType> struct SPMV {
  using VALTYPE = typename CSRMatrixType::VALTYPE;
  void setMatrix(CSRMatrixType const *matrix) {
    _matrix = matrix;
    A_.reset();
  }
  void preprocess() {
    A_ = std::make_shared<detail::MatHandle>(_matrix->rows, _matrix->cols, _matrix->Base(),
                                             _matrix->AI(), _matrix->AJ(), _matrix->AV());
    if (alg() != LA_SPMV_ROW) detail::check(la_spmv_preprocess(A_->h, alg()));
  }
  void operator()(VALTYPE const *b, VALTYPE *x, VALTYPE alpha = 1., VALTYPE beta = 0.) const {
    if (!A_) const_cast<SPMV *>(this)->preprocess();
    else detail::check(la_mat_set_values(A_->h, _matrix->AV()));
    detail::host_spmv(A_->h, alg(), _matrix->rows, _matrix->cols, b, x, alpha, beta);
  }
  // device-resident form: b and x are device pointers
  void apply_device(VALTYPE const *b, VALTYPE *x, VALTYPE alpha = 1., VALTYPE beta = 0.) const {
    if (!A_) const_cast<SPMV *>(this)->preprocess();
    else if (!values_constant_) detail::check(la_mat_set_values(A_->h, _matrix->AV()));
    detail::check(la_spmv(A_->h, alg(), LA_F64, LA_F64, LA_F64, alpha, b, beta, x));
  }
  // the caller promises not to change _matrix's values (apply_device only)
  void values_constant(bool c) { values_constant_ = c; }
  static constexpr la_spmv_alg alg() {
    if constexpr (std::is_same_v<SPMVType, SerialSPMV> || std::is_same_v<SPMVType, ParallelSPMV>)
      return LA_SPMV_ROW;
    else if constexpr (std::is_same_v<SPMVType, SegSumSPMV<>>)
      return LA_SPMV_SEGSUM;
    else
      return LA_SPMV_ALBUS;
  }
  CSRMatrixType const *_matrix = nullptr;
  std::shared_ptr<detail::MatHandle> A_;
  SPMVType _spmv;
  bool values_constant_ = false;
};

// ------------------------------------------------------------- host helpers
// matrix_utils.hpp:637-675
template <TriangularMatrix TS = L, typename VEC>
int TopologicalSort2(int nodes, int base, int const *ai, int const *aj, VEC &iperm,
                     VEC &prefix) {
  iperm.resize(nodes);
  std::vector<int> pre(static_cast<size_t>(nodes) + 2);
  int levels = 0;
  detail::check(la_topo_sort2(nodes, base, ai, aj, TS == U, iperm.data(), pre.data(), &levels));
  prefix.assign(pre.begin(), pre.begin() + levels + 1);
  return levels;
}

// matrix_utils.hpp:720-803
template <typename CSRMatrixType>
void SplitLDU(int rows, int base, int const *ai, int const *aj, double const *av,
              CSRMatrixType &Lm, std::vector<double> &D, CSRMatrixType &Um) {
  Lm.rows = Lm.cols = Um.rows = Um.cols = rows;
  Lm.ai.resize(rows + 1);
  Um.ai.resize(rows + 1);
  int ln = 0, un = 0;
  detail::check(la_split_ldu(rows, base, ai, aj, av, Lm.ai.data(), nullptr, nullptr, nullptr,
                             Um.ai.data(), nullptr, nullptr, &ln, &un));
  Lm.aj.resize(ln); Lm.av.resize(ln); Um.aj.resize(un); Um.av.resize(un); D.resize(rows);
  detail::check(la_split_ldu(rows, base, ai, aj, av, Lm.ai.data(), Lm.aj.data(), Lm.av.data(),
                             D.data(), Um.ai.data(), Um.aj.data(), Um.av.data(), &ln, &un));
}

// matrix_utils.hpp:805-879
template <TriangularMatrix TS = U, typename CSRMatrixType>
void SplitTriangle(int rows, int base, int const *ai, int const *aj, double const *av,
                   CSRMatrixType &T) {
  T.rows = T.cols = rows;
  T.ai.resize(rows + 1);
  int t = 0;
  detail::check(la_split_triangle(rows, base, ai, aj, av, TS == U, T.ai.data(), nullptr, nullptr, &t));
  T.aj.resize(t);
  T.av.resize(t);
  detail::check(la_split_triangle(rows, base, ai, aj, av, TS == U, T.ai.data(), T.aj.data(),
                                  T.av.data(), &t));
}

// matrix_utils.hpp:677-704
inline bool Diagonal(int rows, int base, int const *ai, int const *aj, double const *av,
                     int *diagpos, double *diag) {
  int ok = 0;
  detail::check(la_diagonal(rows, base, ai, aj, av, diagpos, diag, &ok));
  return ok != 0;
}

// matrix_utils.hpp:311-390
inline void ParallelTranspose2(int rows, int cols, int base, int const *ai, int const *aj,
                               double const *av, int *ait, int *ajt, double *avt) {
  detail::check(la_transpose(rows, cols, base, ai, aj, av, ait, ajt, avt));
}

// matrix_utils.hpp:429-472
inline void permute(int rows, int base, int const *ai, int const *aj, double const *av,
                    int const *iperm, int const *perm, int *pai, int *paj, double *pav) {
  detail::check(la_permute(rows, base, ai, aj, av, iperm, perm, pai, paj, pav));
}

// ------------------------------------------------------------- triangular
// ForwardSubstitution (triangle_solve.hpp:32-43): strict L, unit diagonal.
inline void ForwardSubstitution(int size, int base, int const *ai, int const *aj,
                                double const *av, double const *b, double *x) {
  detail::TrsvHandle T(size, base, ai, aj, av, nullptr, LA_LOWER);
  detail::DeviceBuffer db(sizeof(double) * size), dx(sizeof(double) * size);
  db.upload(b, sizeof(double) * size);
  detail::check(la_trsv_solve(T.h, LA_TRSV_LEVEL, LA_F64, LA_F64, db.get(), dx.get()));
  dx.download(x, sizeof(double) * size);
}
// BackwardSubstitution (triangle_solve.hpp:59-71): strict U plus diag.
inline void BackwardSubstitution(int size, int base, int const *ai, int const *aj,
                                 double const *av, double const *diag, double const *b,
                                 double *x) {
  detail::TrsvHandle T(size, base, ai, aj, av, diag, LA_UPPER);
  detail::DeviceBuffer db(sizeof(double) * size), dx(sizeof(double) * size);
  db.upload(b, sizeof(double) * size);
  detail::check(la_trsv_solve(T.h, LA_TRSV_LEVEL, LA_F64, LA_F64, db.get(), dx.get()));
  dx.download(x, sizeof(double) * size);
}
// ForwardSubstitutionT (triangle_solve.hpp:87-98): (I + U^T) x = b from the
// CSR of strict U.
inline void ForwardSubstitutionT(int size, int base, int const *ai, int const *aj,
                                 double const *av, double const *b, double *x) {
  detail::TrsvHandle T(size, base, ai, aj, av, nullptr, LA_UPPER);
  detail::DeviceBuffer db(sizeof(double) * size), dx(sizeof(double) * size);
  db.upload(b, sizeof(double) * size);
  detail::check(la_trsv_solve_transpose(T.h, LA_TRSV_LEVEL, LA_F64, LA_F64, db.get(), dx.get()));
  dx.download(x, sizeof(double) * size);
}
// LevelSchedule{Forward,Backward}Substitution (triangle_solve.hpp:100-145):
// the level sets are recomputed by the plan (identical to the ones passed).
inline void LevelScheduleForwardSubstitution(int const *, int const *, int, int rows, int base,
                                             int const *ai, int const *aj, double const *av,
                                             double const *b, double *x) {
  ForwardSubstitution(rows, base, ai, aj, av, b, x);
}
inline void LevelScheduleBackwardSubstitution(int const *, int const *, int, int rows, int base,
                                              int const *ai, int const *aj, double const *av,
                                              double const *diag, double const *b, double *x) {
  BackwardSubstitution(rows, base, ai, aj, av, diag, b, x);
}

// OptimizedTriangularSolve (triangle_solve.hpp:147-208): analysis builds the
// device plan (level sets + level-ordered copy); Barrier runs one launch per
// level, NoBarrier / NoBarrierSuperNode the sync-free kernel. Unlike the
// reference, diag is copied at analysis (the reference keeps the pointer,
// triangle_solve.tpp:22).
template <FBSubstitutionType FBST = FBSubstitutionType::Barrier,
          TriangularMatrix TS = TriangularMatrix::L, typename ROWTYPE = int,
          typename COLTYPE = int, typename VALTYPE = double>
class OptimizedTriangularSolve {
 public:
  explicit OptimizedTriangularSolve(int num_threads = 1) : _nthreads(num_threads) {}
  void analysis(COLTYPE rows, int base, ROWTYPE const *ai, COLTYPE const *aj,
                VALTYPE const *av, VALTYPE const *diag = nullptr) {
    _size = rows;
    T_ = std::make_shared<detail::TrsvHandle>(rows, base, ai, aj, av, diag,
                                              TS == U ? LA_UPPER : LA_LOWER);
    b_.resize(sizeof(double) * (rows ? rows : 1));
    x_.resize(sizeof(double) * (rows ? rows : 1));
  }
  void operator()(VALTYPE const *b, VALTYPE *x) const {
    b_.upload(b, sizeof(double) * _size);
    apply_device(static_cast<const VALTYPE *>(b_.get()), static_cast<VALTYPE *>(x_.get()));
    x_.download(x, sizeof(double) * _size);
  }
  void apply_device(VALTYPE const *b, VALTYPE *x) const {
    detail::check(la_trsv_solve(T_->h, mode(), LA_F64, LA_F64, b, x));
  }
  int get_num_threads() const { return _nthreads; }
  int levels() const {
    int n = 0, lv = 0;
    int64_t nnz = 0;
    detail::check(la_trsv_info(T_->h, &n, &lv, &nnz));
    return lv;
  }

 private:
  static constexpr la_trsv_mode mode() {
    return FBST == FBSubstitutionType::Barrier ? LA_TRSV_LEVEL : LA_TRSV_SYNCFREE;
  }
  int _nthreads;
  COLTYPE _size = 0;
  std::shared_ptr<detail::TrsvHandle> T_;
  mutable detail::DeviceBuffer b_, x_;
};

// ------------------------------------------------------------- IC(k)
// ICCLevelSymbolic0..3 (precond.hpp:39-613) produce the same pattern; all
// four names build it the same way.
template <typename CSRMatrixType>
void ICCLevelSymbolic0(int size, int base, int const *ai, int const *aj, int const *diag_pos,
                       int lvl, CSRMatrixType &icc) {
  icc.rows = icc.cols = size;
  icc.ai.resize(size + 1);
  int64_t nnz = 0;
  detail::check(la_icc_symbolic(size, base, ai, aj, diag_pos, lvl, icc.ai.data(), nullptr, &nnz));
  icc.aj.resize(nnz);
  detail::check(la_icc_symbolic(size, base, ai, aj, diag_pos, lvl, icc.ai.data(), icc.aj.data(), &nnz));
  icc.av.resize(nnz);
}
template <typename M>
void ICCLevelSymbolic1(int s, int b, int const *ai, int const *aj, int const *d, int l, M &m) {
  ICCLevelSymbolic0(s, b, ai, aj, d, l, m);
}
template <typename M>
void ICCLevelSymbolic2(int s, int b, int const *ai, int const *aj, int const *d, int l, M &m) {
  ICCLevelSymbolic0(s, b, ai, aj, d, l, m);
}
template <typename M>
void ICCLevelSymbolic3(int s, int b, int const *ai, int const *aj, int const *d, int l, M &m) {
  ICCLevelSymbolic0(s, b, ai, aj, d, l, m);
}
// precond.hpp:615-693
inline bool ICCLevelNumeric(int size, int base, int const *ai, int const *aj, double const *av,
                            int const *diag_pos, int lvl, double omega, int const *icc_ai,
                            int const *icc_aj, double *icc_av) {
  (void)lvl;
  int ok = 0;
  detail::check(la_icc_numeric(size, base, ai, aj, av, diag_pos, omega, icc_ai, icc_aj, icc_av, &ok));
  return ok != 0;
}
}  // namespace matrix_utils

// =====================================================================
namespace mkl_wrapper {

// Owning CSR matrix (mkl_wrapper/mkl_sparse_mat.h:9-123, container part):
// host arrays in the caller's base, uploaded to the device on first use.
class mkl_sparse_mat {
 public:
  mkl_sparse_mat() = default;
  mkl_sparse_mat(int row, int col, const std::vector<int> &ai, const std::vector<int> &aj,
                 const std::vector<double> &av)
      : _nrow(row), _ncol(col), _base(ai.empty() ? 0 : ai[0]),
        _ai(new int[ai.size()]), _aj(new int[aj.size()]), _av(new double[av.size()]) {
    std::copy(ai.begin(), ai.end(), _ai.get());
    std::copy(aj.begin(), aj.end(), _aj.get());
    std::copy(av.begin(), av.end(), _av.get());
    _nnz = static_cast<int>(ai[row] - ai[0]);
  }
  mkl_sparse_mat(int row, int col, std::shared_ptr<int[]> ai, std::shared_ptr<int[]> aj,
                 std::shared_ptr<double[]> av)
      : _nrow(row), _ncol(col), _base(ai[0]), _ai(ai), _aj(aj), _av(av),
        _nnz(static_cast<int>(ai[row] - ai[0])) {}
  virtual ~mkl_sparse_mat() = default;

  int rows() const { return _nrow; }
  int cols() const { return _ncol; }
  int nnz() const { return _nnz; }
  int mkl_base() const { return _base; }
  std::shared_ptr<int[]> get_ai() const { return _ai; }
  std::shared_ptr<int[]> get_aj() const { return _aj; }
  std::shared_ptr<double[]> get_av() const { return _av; }

  // mkl_sparse_mat::mult_vec (mkl_sparse_mat.cpp:325-329): x = A b
  virtual bool mult_vec(double const *const b, double *const x) {
    detail::host_spmv(device(), LA_SPMV_AUTO, _nrow, _ncol, b, x, 1.0, 0.0);
    return true;
  }
  // mkl_sparse_mat::transpose_mult_vec (mkl_sparse_mat.cpp:331-336): x = A^T b
  virtual bool transpose_mult_vec(double const *const b, double *const x) {
    detail::DeviceBuffer db(sizeof(double) * _nrow), dx(sizeof(double) * _ncol);
    db.upload(b, sizeof(double) * _nrow);
    detail::check(la_spmv_transpose(device(), LA_F64, LA_F64, 1.0, db.get(), 0.0, dx.get()));
    dx.download(x, sizeof(double) * _ncol);
    return true;
  }
  // mkl_sparse_mat::rowwiseSqrtNorm (mkl_sparse_mat.cpp:517-533)
  std::vector<double> rowwiseSqrtNorm() const {
    std::vector<double> d(_nrow);
    detail::check(la_rowwise_sqrt_norm(_nrow, _base, _ai.get(), _av.get(), d.data()));
    return d;
  }
  // mkl_sparse_mat::DtAD (mkl_sparse_mat.cpp:535-544): a_ij *= d_i d_j, in
  // place; the device copy is dropped and re-uploaded on next use
  void DtAD(const std::vector<double> &diag) {
    detail::check(la_dtad(_nrow, _base, _ai.get(), _aj.get(), _av.get(), diag.data()));
    _dev.reset();
  }
  // mkl_sparse_mat::to_one_based / to_zero_based (mkl_sparse_mat.cpp:235-267):
  // shift the index arrays in place; the device copy is rebuilt on next use
  void to_one_based() { rebase(1); }
  void to_zero_based() { rebase(0); }
  // preconditioners override this (z = M^-1 r)
  virtual bool solve(double const *const, double *const) { return false; }
  virtual bool solve_device(double const *const, double *const) { return false; }
  virtual la_prec_t prec_handle() const { return nullptr; }

  la_mat_t device() {
    if (!_dev) _dev = std::make_shared<detail::MatHandle>(_nrow, _ncol, _base, _ai.get(), _aj.get(), _av.get());
    return _dev->h;
  }

 protected:
  void rebase(int b) {
    if (_base == b || !_ai) return;
    const int d = b - _base;
    for (int i = 0; i < _nnz; ++i) _aj[i] += d;
    for (int i = 0; i <= _nrow; ++i) _ai[i] += d;
    _base = b;
    _dev.reset();
  }
  int _nrow = 0, _ncol = 0, _base = 0;
  std::shared_ptr<int[]> _ai, _aj;
  std::shared_ptr<double[]> _av;
  int _nnz = 0;
  std::shared_ptr<detail::MatHandle> _dev;
};

// Symmetric matrix given by its full pattern, stored as the upper triangle
// (mkl_sparse_mat_sym, mkl_sparse_mat.cpp:669-694). Products use the full
// matrix (reassembled once for the device).
class mkl_sparse_mat_sym : public mkl_sparse_mat {
 public:
  explicit mkl_sparse_mat_sym(const mkl_sparse_mat &A) : full_(A) {
    matrix_utils::CSRMatrixVec<int, int, double> Uv;
    matrix_utils::SplitTriangle<matrix_utils::U>(A.rows(), A.mkl_base(), A.get_ai().get(),
                                                 A.get_aj().get(), A.get_av().get(), Uv);
    _nrow = _ncol = A.rows();
    _base = A.mkl_base();
    _ai.reset(new int[Uv.ai.size()]);
    _aj.reset(new int[Uv.aj.size()]);
    _av.reset(new double[Uv.av.size()]);
    std::copy(Uv.ai.begin(), Uv.ai.end(), _ai.get());
    std::copy(Uv.aj.begin(), Uv.aj.end(), _aj.get());
    std::copy(Uv.av.begin(), Uv.av.end(), _av.get());
    _nnz = static_cast<int>(Uv.aj.size());
  }
  bool mult_vec(double const *const b, double *const x) override { return full_.mult_vec(b, x); }
  la_mat_t full_device() { return full_.device(); }

 private:
  mkl_sparse_mat full_;
};

// common part of the ILU / IC preconditioners: the factor and its apply
class device_preconditioner : public mkl_sparse_mat {
 public:
  bool solve(double const *const b, double *const x) override {
    detail::DeviceBuffer db(sizeof(double) * _nrow), dx(sizeof(double) * _nrow);
    db.upload(b, sizeof(double) * _nrow);
    solve_device(static_cast<const double *>(db.get()), static_cast<double *>(dx.get()));
    dx.download(x, sizeof(double) * _nrow);
    return true;
  }
  bool solve_device(double const *const b, double *const x) override {
    detail::check(la_prec_apply(P_.get(), trsv_mode_, LA_F64, LA_F64, b, x));
    return true;
  }
  la_prec_t prec_handle() const override { return P_.get(); }
  void set_trsv_mode(la_trsv_mode m) { trsv_mode_ = m; }
  la_trsv_mode trsv_mode() const { return trsv_mode_; }
  // sweeps per triangle of LA_TRSV_JACOBI applies (no counterpart in the reference)
  void set_sweeps(int k) {
    detail::check(la_prec_set_sweeps(P_.get(), k));
    sweeps_ = k;
  }

 protected:
  void upload_factor(bool ic) {
    la_prec_t p = nullptr;
    if (ic) detail::check(la_prec_create_ic(_nrow, _base, _ai.get(), _aj.get(), _av.get(), &p));
    else detail::check(la_prec_create_ilu(_nrow, _base, _ai.get(), _aj.get(), _av.get(), &p));
    P_ = std::shared_ptr<la_prec_s>(p, [](la_prec_t q) { if (q) la_prec_destroy(q); });
    if (sweeps_ > 0) detail::check(la_prec_set_sweeps(p, sweeps_));
  }
  std::shared_ptr<la_prec_s> P_;
  int sweeps_ = 0;  // 0: the library's default
  la_trsv_mode trsv_mode_ = LA_TRSV_SYNCFREE;
};

// incomplete_lu_k (mkl_wrapper/incomplete_lu.h:55-70, .cpp:153-325)
class incomplete_lu_k : public device_preconditioner {
 public:
  void set_level(int level) { _level = level; }
  bool symbolic_factorize(mkl_sparse_mat const *const A) {
    _nrow = A->rows();
    _ncol = A->cols();
    _base = A->mkl_base();
    std::vector<int> ai(_nrow + 1);
    first_upper_.resize(_nrow);
    int64_t nnz = 0;
    detail::check(la_iluk_symbolic(_nrow, _base, A->get_ai().get(), A->get_aj().get(), _level,
                                   ai.data(), nullptr, first_upper_.data(), &nnz));
    _ai.reset(new int[_nrow + 1]);
    _aj.reset(new int[nnz]);
    _av.reset(new double[nnz]);
    detail::check(la_iluk_symbolic(_nrow, _base, A->get_ai().get(), A->get_aj().get(), _level,
                                   _ai.get(), _aj.get(), first_upper_.data(), &nnz));
    _nnz = static_cast<int>(nnz);
    return true;
  }
  // The factor is left 1-based like the reference's (incomplete_lu.cpp:320-323);
  // a later call first brings the pattern back to A's base.
  bool numeric_factorize(mkl_sparse_mat const *const A) {
    rebase(A->mkl_base());
    const int st = la_iluk_numeric(_nrow, _base, A->get_ai().get(), A->get_aj().get(),
                                   A->get_av().get(), _ai.get(), _aj.get(), first_upper_.data(),
                                   _av.get());
    if (st == LA_ERR_NUMERIC) {
      std::fprintf(stderr, "no element on diagonal!\n");
      return false;
    }
    detail::check(st);
    upload_factor(false);
    to_one_based();
    return true;
  }

 private:
  int _level = 0;
  std::vector<int> first_upper_;
};

// Both IC(k) classes leave the factor 1-based like the reference's
// (incomplete_cholesky.cpp:276-279, :337-340). Symbolic phase shared by both: the level-of-fill pattern of
// the upper triangle of A (incomplete_cholesky.cpp:286-313 via
// ICCLevelSymbolic; incomplete_cholesky_k's own symbolic phase at :74-193
// yields the same pattern, tests/precond_test.cpp:63-116).
class incomplete_cholesky_common : public device_preconditioner {
 public:
  void set_level(int level) { _level = level; }
  void shift(bool s) { _shift = s; }
  bool symbolic_factorize(mkl_sparse_mat const *const A) {
    _nrow = _ncol = A->rows();
    _base = A->mkl_base();
    diag_pos_.resize(_nrow);
    int ok = 0;
    detail::check(la_diagonal(_nrow, _base, A->get_ai().get(), A->get_aj().get(), nullptr,
                              diag_pos_.data(), nullptr, &ok));
    if (!ok) {
      std::fprintf(stderr, "Failed to get diagonal\n");
      return false;
    }
    std::vector<int> ai(_nrow + 1);
    int64_t nnz = 0;
    detail::check(la_icc_symbolic(_nrow, _base, A->get_ai().get(), A->get_aj().get(),
                                  diag_pos_.data(), _level, ai.data(), nullptr, &nnz));
    _ai.reset(new int[_nrow + 1]);
    _aj.reset(new int[nnz]);
    _av.reset(new double[nnz]);
    detail::check(la_icc_symbolic(_nrow, _base, A->get_ai().get(), A->get_aj().get(),
                                  diag_pos_.data(), _level, _ai.get(), _aj.get(), &nnz));
    _nnz = static_cast<int>(nnz);
    return true;
  }

 protected:
  int _level = 0;
  bool _shift = false;
  double _initial_shift = 1e-3;
  int _nrestart = 20;
  std::vector<int> diag_pos_;
};

// incomplete_cholesky_k_2 (incomplete_cholesky.cpp:286-342): numeric phase
// with a multiplicative diagonal shift a_ii (1 + s), s doubled from 1e-3 up
// to 20 times when shift(true); fails on a negative pivot.
class incomplete_cholesky_k_2 : public incomplete_cholesky_common {
 public:
  bool numeric_factorize(mkl_sparse_mat const *const A) {
    rebase(A->mkl_base());
    double shift = 0.0;
    for (int iter = 0;; ) {
      int ok = 0;
      detail::check(la_icc_numeric(_nrow, _base, A->get_ai().get(), A->get_aj().get(),
                                   A->get_av().get(), diag_pos_.data(), shift, _ai.get(),
                                   _aj.get(), _av.get(), &ok));
      if (ok) break;
      if (!_shift || ++iter > _nrestart) return false;
      shift = std::max(2 * shift, _initial_shift);
      std::printf("shift: %g\n", shift);
    }
    upload_factor(true);
    to_one_based();
    return true;
  }
};

// incomplete_cholesky_k (incomplete_cholesky.cpp:195-274): numeric phase
// with an ADDITIVE diagonal shift a_ii + s, fails on a pivot <= 0. With
// shift(true) the first s is 1e-3 - min a_ii when min a_ii <= 0 (else 0),
// then s = max(1e-3, 2 s) on each failure, at most 20 retries.
class incomplete_cholesky_k : public incomplete_cholesky_common {
 public:
  bool numeric_factorize(mkl_sparse_mat const *const A) {
    rebase(A->mkl_base());
    const double *av = A->get_av().get();
    double shift = 0.0;
    if (_shift) {
      double min_diag = std::numeric_limits<double>::max();
      for (int i = 0; i < _nrow; ++i) min_diag = std::min(min_diag, av[diag_pos_[i] - _base]);
      if (min_diag <= 0.0) shift = _initial_shift - min_diag;
      std::printf("init shift: %g\n", shift);
    }
    for (int iter = 0;; ) {
      int ok = 0;
      detail::check(la_ick_numeric(_nrow, _base, A->get_ai().get(), A->get_aj().get(), av,
                                   diag_pos_.data(), shift, _ai.get(), _aj.get(), _av.get(),
                                   &ok));
      if (ok) break;
      if (!_shift || ++iter > _nrestart) return false;
      shift = std::max(_initial_shift, 2.0 * shift);
      std::printf("shift: %g\n", shift);
    }
    upload_factor(true);
    to_one_based();
    return true;
  }
};

// mkl_wrapper/mkl_solver.h:10-87
class mkl_solver {
 public:
  virtual ~mkl_solver() = default;
  virtual bool solve(double const *const b, double *const x) = 0;
  void set_print_level(int pl) { _print_level = pl; }

 protected:
  int _print_level = 0;
};

class mkl_iterative_solver : public mkl_solver {
 public:
  void set_max_iters(int n) { _maxiter = n; }
  void set_rel_tol(double tol) { _rel_tol = tol; }
  void set_abs_tol(double tol) { _abs_tol = tol; }
  // extensions with no counterpart in the reference
  void set_precision(la_precision p) { _precision = p; }
  void set_stop_rule(la_stop_rule r) { _stop = r; }
  void set_inner_tol(double t) { _inner_tol = t; }
  void set_reliable_delta(double d) { _reliable_delta = d; }
  const la_solver_result &last_result() const { return _res; }

 protected:
  la_solver_opts opts(la_prec_t P, const mkl_sparse_mat *prec) const {
    la_solver_opts o;
    la_solver_opts_default(&o);
    o.rel_tol = _rel_tol;
    o.abs_tol = _abs_tol;
    o.max_iters = _maxiter;
    o.precision = _precision;
    o.stop_rule = _stop;
    o.inner_tol = _inner_tol;
    o.reliable_delta = _reliable_delta;
    o.print_level = _print_level;
    if (auto dp = dynamic_cast<const device_preconditioner *>(prec)) o.trsv_mode = dp->trsv_mode();
    (void)P;
    return o;
  }
  int _maxiter = 1000;
  double _rel_tol = 1e-8;
  double _abs_tol = 0.;
  la_precision _precision = LA_PREC_FP64;
  la_stop_rule _stop = LA_STOP_TRUE_RESIDUAL;
  double _inner_tol = 1e-5;
  double _reliable_delta = 0.1;
  la_solver_result _res{};
};

namespace detail_solver {
inline la_mat_t device_matrix(mkl_sparse_mat *A) {
  if (auto s = dynamic_cast<mkl_sparse_mat_sym *>(A)) return s->full_device();
  return A->device();
}
inline la_prec_t prec_of(const std::shared_ptr<mkl_sparse_mat> &P) {
  if (!P) return nullptr;
  la_prec_t h = P->prec_handle();
  if (!h) throw error(LA_ERR_ARG, "laamd: preconditioner must be an incomplete_lu_k / incomplete_cholesky_k");
  return h;
}
template <typename F>
bool host_solve(int n, double const *b, double *x, la_solver_result &res, F &&call) {
  laamd::detail::DeviceBuffer db(sizeof(double) * n), dx(sizeof(double) * n);
  db.upload(b, sizeof(double) * n);
  dx.upload(x, sizeof(double) * n);
  const int st = call(static_cast<const double *>(db.get()), static_cast<double *>(dx.get()), &res);
  if (st != LA_OK && st != LA_ERR_NOT_CONVERGED) laamd::detail::check(st);
  dx.download(x, sizeof(double) * n);
  return st == LA_OK;
}
}  // namespace detail_solver

// mkl_pcg_solver (mkl_solver.h:58-70, mkl_solver.cpp:150-251)
class mkl_pcg_solver : public mkl_iterative_solver {
 public:
  mkl_pcg_solver(mkl_sparse_mat *A, std::shared_ptr<mkl_sparse_mat> P = nullptr) : _A(A), _P(P) {}
  bool solve(double const *const b, double *const x) override {
    if (!_A) return false;
    return detail_solver::host_solve(_A->rows(), b, x, _res, [&](const double *db, double *dx,
                                                                  la_solver_result *r) {
      return solve_device_status(db, dx, r);
    });
  }
  // device-resident b and x (x holds x0)
  bool solve_device(double const *const b, double *const x) {
    return solve_device_status(b, x, &_res) == LA_OK;
  }
  mkl_sparse_mat *getPrecond() { return _P.get(); }

 private:
  int solve_device_status(const double *b, double *x, la_solver_result *r) {
    la_prec_t P = detail_solver::prec_of(_P);
    const la_solver_opts o = opts(P, _P.get());
    return la_pcg_solve(detail_solver::device_matrix(_A), P, &o, b, x, r);
  }
  mkl_sparse_mat *_A;
  std::shared_ptr<mkl_sparse_mat> _P;
};

// mkl_fgmres_solver (mkl_solver.h:72-87, mkl_solver.cpp:253-386), with the
// reference's optional right operator _R.
class mkl_fgmres_solver : public mkl_iterative_solver {
 public:
  // R: right operator (mkl_solver.cpp:319-324): the Krylov operator is
  // A R^-1 and x returns u with A R^-1 u = b, as in the reference
  mkl_fgmres_solver(mkl_sparse_mat *A, std::shared_ptr<mkl_sparse_mat> P = nullptr,
                    std::shared_ptr<mkl_sparse_mat> R = nullptr)
      : _A(A), _P(P), _R(R) {}
  void set_restart_steps(int n) { _num_restart = n; }
  bool solve(double const *const b, double *const x) override {
    if (!_A) return false;
    return detail_solver::host_solve(_A->rows(), b, x, _res, [&](const double *db, double *dx,
                                                                  la_solver_result *r) {
      return solve_device_status(db, dx, r);
    });
  }
  bool solve_device(double const *const b, double *const x) {
    return solve_device_status(b, x, &_res) == LA_OK;
  }

 private:
  int solve_device_status(const double *b, double *x, la_solver_result *r) {
    la_prec_t P = detail_solver::prec_of(_P);
    la_solver_opts o = opts(P, _P.get());
    o.restart = _num_restart;  // 0: unrestarted up to the iteration limit, as the reference
    if (_R)
      return la_fgmres_solve_right(detail_solver::device_matrix(_A), P,
                                   detail_solver::prec_of(_R), &o, b, x, r);
    return la_fgmres_solve(detail_solver::device_matrix(_A), P, &o, b, x, r);
  }
  mkl_sparse_mat *_A;
  std::shared_ptr<mkl_sparse_mat> _P, _R;
  int _num_restart = 0;
};

// solver_factory (mkl_solver.h:89-101, mkl_solver.cpp:388-416): name ->
// creator registry; "gmres" is pre-registered with rel 1e-12, abs 1e-15.
class solver_factory {
 public:
  using solver_ptr = std::unique_ptr<mkl_solver>;
  using create_method = std::function<solver_ptr(mkl_sparse_mat &)>;
  solver_factory() {
    _methods["gmres"] = [](mkl_sparse_mat &m) {
      auto s = std::make_unique<mkl_fgmres_solver>(&m);
      s->set_rel_tol(1e-12);
      s->set_abs_tol(1e-15);
      return solver_ptr(std::move(s));
    };
  }
  bool reg(const std::string &name, create_method func) {
    if (_methods.count(name)) return false;
    _methods[name] = std::move(func);
    return true;
  }
  solver_ptr create(const std::string &name, mkl_sparse_mat &m) {
    auto it = _methods.find(name);
    return it == _methods.end() ? nullptr : it->second(m);
  }

 private:
  std::map<std::string, create_method> _methods;
};
}  // namespace mkl_wrapper

namespace utils {
// utils/utils.h:148-162
template <typename T> class singleton {
 public:
  static T &instance() {
    static T inst;
    return inst;
  }
};
}  // namespace utils
}  // namespace laamd
