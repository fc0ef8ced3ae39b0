// Preconditioned Krylov loops driving the device kernels:
//  * PCG (the loop of mkl_pcg_solver::solve, mkl_wrapper/mkl_solver.cpp:
//    150-251): textbook PCG (Saad, Alg. 9.1); the reference's stopping test
//    is the true residual ||b - A x|| / ||b|| (or < abs_tol) evaluated every
//    iteration with an extra SpMV (:211-225).
//  * FGMRES(m) (mkl_fgmres_solver::solve, :253-386): right/flexible
//    preconditioning, classical Gram-Schmidt with one re-orthogonalisation
//    pass (one multi-dot + one multi-axpy per pass), Givens rotations on
//    the host; stops on the Arnoldi residual estimate / ||b|| (:335-346) or
//    a Hessenberg breakdown h_{k+1,k} < 1e-16 (:353-367).
// Mixed precision: the iteration runs with FP32 vectors and FP32/BF16
// matrix and factors (FP32 accumulation, FP64 dot products); x and the
// true residual r = b - A x (FP64 matrix) stay in FP64.
//  * PCG, reliable_delta > 0: reliable updates (Sleijpen & van der Vorst,
//    "Reliable updated residuals in hybrid Bi-CG methods", Computing 56,
//    1996; Clark et al., "Solving lattice QCD systems of equations using
//    mixed precision solvers on GPUs", 2010): one PCG iteration on b/||b||
//    whose FP32 increment of x is folded into the FP64 x, and whose
//    residual is replaced by the FP64 true residual, each time the
//    recurrence residual falls below delta times its largest value since
//    the last replacement, and before stopping. The search direction is
//    kept across a replacement, so the Krylov space is not thrown away.
//  * otherwise (and FGMRES): restarted iterative refinement -- solve
//    A d = r / ||r|| approximately, x += ||r|| d in FP64, repeat.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <vector>

#include "device.hpp"

namespace laamd {

namespace {
struct Work {
  int64_t n = 0;
  int m = 0;
  DevBuf v64[7], v32[7];
  DevBuf V, Z;  // GMRES bases
  la_dtype basis_dt = LA_F64;
  // GMRES Hessenberg column (two Gram-Schmidt passes and the norm) on the
  // device, and its pinned host staging; sized for the restart length
  DevBuf hcol;
  double *hpin = nullptr;
  int hcap = 0;
  Scratch sc;
  void ensure_hcol(int cap) {
    if (cap <= hcap) return;
    hcol.alloc(sizeof(double) * (2 * static_cast<size_t>(cap) + 1));
    if (hpin) (void)hipHostFree(hpin);
    hpin = nullptr;
    LA_HIP(hipHostMalloc(reinterpret_cast<void **>(&hpin),
                         sizeof(double) * (2 * static_cast<size_t>(cap) + 1),
                         hipHostMallocDefault));
    hcap = cap;
  }
  Work() = default;
  Work(const Work &) = delete;
  Work &operator=(const Work &) = delete;
  ~Work() {
    if (hpin) (void)hipHostFree(hpin);
  }
};

thread_local std::unique_ptr<Work> tl_work;

Work &work_for(int64_t n) {
  if (!tl_work || tl_work->n != n) {
    tl_work = std::make_unique<Work>();
    tl_work->n = n;
    for (int i = 0; i < 7; ++i) {
      tl_work->v64[i].alloc(sizeof(double) * std::max<int64_t>(n, 1));
      tl_work->v32[i].alloc(sizeof(float) * std::max<int64_t>(n, 1));
    }
  }
  return *tl_work;
}

double fetch(Scratch &S, const double *dev, hipStream_t s) {
  LA_HIP(hipMemcpyAsync(S.host, dev, sizeof(double), hipMemcpyDeviceToHost, s));
  LA_HIP(hipStreamSynchronize(s));
  return S.host[0];
}

double norm2(int64_t n, la_dtype dt, const void *x, Scratch &S, hipStream_t s) {
  double *d = S.scalars.as<double>() + 200;
  dot_dev(n, dt, x, x, S.partials.as<double>(), d, s);
  return std::sqrt(fetch(S, d, s));
}

// value types of the inner solve: the matrix, and the preconditioner factors
la_dtype inner_val_dtype(int precision) {
  return precision == LA_PREC_MIXED_BF16 || precision == LA_PREC_MIXED_BF16_F32P ? LA_BF16
                                                                                 : LA_F32;
}
la_dtype inner_prec_dtype(int precision) {
  return precision == LA_PREC_MIXED_BF16 || precision == LA_PREC_MIXED_F32_BF16P ? LA_BF16
                                                                                 : LA_F32;
}

// z = M^-1 r, or M^-T r when the solve is with A^T
thread_local bool tl_transpose = false;
// right operator of FGMRES (mkl_fgmres_solver's _R, mkl_solver.cpp:319-324):
// the Krylov operator is A R^-1
thread_local Prec *tl_right = nullptr;
void apply_M(Prec *P, const la_solver_opts &o, la_dtype val, la_dtype vec, const void *r,
             void *z, int64_t n, hipStream_t s) {
  const la_trsv_mode m = static_cast<la_trsv_mode>(o.trsv_mode);
  if (val != LA_F64) val = inner_prec_dtype(o.precision);
  if (!P) copy_vec(n, vec, r, vec, z, 1.0, s);
  else if (tl_transpose) prec_apply_transpose(*P, m, val, vec, r, z, s);
  else prec_apply(*P, m, val, vec, r, z, s);
}

// y = alpha * Op x + beta * z with Op = A, or A R^-1 under a right operator
// (R^-1 x goes through the spare work vector 6)
void op_apply(Mat &A, const la_solver_opts &o, la_spmv_alg alg, la_dtype val, la_dtype vec,
              la_dtype acc, double alpha, const void *x, double beta, const void *z, void *y,
              Work &W, hipStream_t s) {
  if (!tl_right) {
    spmv(A, alg, val, vec, acc, alpha, x, beta, z, y, s);
    return;
  }
  void *t = (vec == LA_F64 ? W.v64 : W.v32)[6].p;
  prec_apply(*tl_right, static_cast<la_trsv_mode>(o.trsv_mode), val, vec, x, t, s);
  spmv(A, alg, val, vec, acc, alpha, t, beta, z, y, s);
}

// One PCG run on (val, vec) precision. x holds x0; vectors b/x in `vec`.
// In `true_residual` mode the stop test uses ||b - A x|| (FP64 vectors
// only); otherwise the recurrence residual, relative to ||b||.
struct CgResult {
  bool converged = false;
  int iters = 0;
  double last_res = 0;
};

CgResult pcg_core(Mat &A, Prec *P, const la_solver_opts &o, la_dtype val, la_dtype vec,
                  const void *b, void *x, double nb, double rel_tol, double abs_tol,
                  int maxiter, bool true_residual, Work &W, hipStream_t s) {
  const int64_t n = A.rows;
  const la_dtype acc = (val == LA_F64 && vec == LA_F64) ? LA_F64 : LA_F32;
  DevBuf *vb = vec == LA_F64 ? W.v64 : W.v32;
  void *r = vb[0].p, *z = vb[1].p, *p = vb[2].p, *q = vb[3].p, *t = vb[4].p;
  double *sc = W.sc.scalars.as<double>(), *part = W.sc.partials.as<double>();
  const la_spmv_alg alg = static_cast<la_spmv_alg>(o.spmv_alg);
  CgResult R;
  // partial-sum slots: <p,q>, <r,r>, <t,t>, and <r,z> double-buffered
  double *p_pq = part, *p_rr = part + kRedBlocks, *p_tt = part + 2 * kRedBlocks;
  double *p_rz[2] = {part + 3 * kRedBlocks, part + 4 * kRedBlocks};
  int cur = 0;
  (void)sc;
  // r = b - A x ; z = M r ; p = z ; rz = <r,z>
  spmv(A, alg, val, vec, acc, -1.0, x, 1.0, b, r, s);
  apply_M(P, o, val, vec, r, z, n, s);
  copy_vec(n, vec, z, vec, p, 1.0, s);
  dot_partials(n, vec, r, z, p_rz[cur], s);
  for (int it = 1; it <= maxiter; ++it) {
    spmv(A, alg, val, vec, acc, 1.0, p, 0.0, q, q, s);
    dot_partials(n, vec, p, q, p_pq, s);
    cg_update_xr(n, vec, x, r, p, q, p_rz[cur], p_pq, p_rr, s);
    const double *p_check = p_rr;
    if (true_residual) {
      spmv(A, alg, val, vec, acc, -1.0, x, 1.0, b, t, s);
      dot_partials(n, vec, t, t, p_tt, s);
      p_check = p_tt;
    }
    const uint64_t tk = report_sum(W.sc, p_check, s);
    // the next direction is queued before the stop test is read, so the
    // device runs it while the host waits; a converged run discards it
    // (x and r are final before it)
    apply_M(P, o, val, vec, r, z, n, s);
    dot_partials(n, vec, r, z, p_rz[cur ^ 1], s);
    cg_update_p(n, vec, p, z, p_rz[cur ^ 1], p_rz[cur], s);
    const double res = std::sqrt(wait_report(W.sc, tk, s));
    R.iters = it;
    R.last_res = res;
    if (o.print_level == 1) std::fprintf(stderr, "%3d = %lg/%lg=%lg\n", it, res, nb, res / nb);
    if (res / nb < rel_tol || res < abs_tol || !std::isfinite(res)) {
      R.converged = std::isfinite(res);
      break;
    }
    cur ^= 1;
  }
  LA_HIP(hipStreamSynchronize(s));
  return R;
}

// One FGMRES run (restart cycles until converged or maxiter) on (val, vec).
CgResult fgmres_core(Mat &A, Prec *P, const la_solver_opts &o, la_dtype val, la_dtype vec,
                     const void *b, void *x, double nb, double rel_tol, double abs_tol,
                     int maxiter, int m, Work &W, hipStream_t s) {
  const int64_t n = A.rows;
  const la_dtype acc = (val == LA_F64 && vec == LA_F64) ? LA_F64 : LA_F32;
  const size_t es = dtype_size(vec);
  if (W.m < m || W.basis_dt != vec || !W.V.p) {
    W.V.alloc(es * static_cast<size_t>(n) * (m + 1));
    W.Z.alloc(es * static_cast<size_t>(n) * m);
    W.m = m;
    W.basis_dt = vec;
  }
  char *V = static_cast<char *>(W.V.p), *Z = static_cast<char *>(W.Z.p);
  auto Vj = [&](int j) { return static_cast<void *>(V + es * static_cast<size_t>(n) * j); };
  auto Zj = [&](int j) { return static_cast<void *>(Z + es * static_cast<size_t>(n) * j); };
  DevBuf *vb = vec == LA_F64 ? W.v64 : W.v32;
  void *w = vb[0].p;
  double *part = W.sc.partials.as<double>();
  // Hessenberg column: h1 and h2 hold the two passes' m + 1 coefficients,
  // hn the norm; the update coefficients reuse h1
  W.ensure_hcol(m + 1);
  const int cap = m + 1;
  double *h1 = W.hcol.as<double>(), *h2 = h1 + cap, *hn = h2 + cap;
  const la_spmv_alg alg = static_cast<la_spmv_alg>(o.spmv_alg);
  std::vector<double> H(static_cast<size_t>(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  // pinned staging for the Hessenberg column and the update coefficients
  // (a pageable buffer would add a staged copy per iteration)
  double *hbuf = W.hpin;
  // Gram-Schmidt against basis vectors [0, k): the dot products and the
  // update in groups of at most kMaxDots vectors (one group up to m = 63)
  auto gs_dots = [&](int k, double *out) {
    for (int c0 = 0; c0 < k; c0 += kMaxDots)
      multidot_dev(n, std::min(kMaxDots, k - c0), vec, Vj(c0), n, w, part, out + c0, s);
  };
  auto gs_update = [&](char *B, int k, const double *h, void *dst) {
    for (int c0 = 0; c0 < k; c0 += kMaxDots)
      multiaxpy_dev(n, std::min(kMaxDots, k - c0), vec, B + es * static_cast<size_t>(n) * c0, n,
                    h + c0, dst, s);
  };
  CgResult R;
  int total = 0;
  bool done = false;
  while (!done) {
    op_apply(A, o, alg, val, vec, acc, -1.0, x, 1.0, b, w, W, s);
    const double beta = norm2(n, vec, w, W.sc, s);
    if (beta == 0.0 || beta / nb < rel_tol || beta < abs_tol) {
      R.converged = true;
      R.last_res = beta;
      break;
    }
    copy_vec(n, vec, w, vec, Vj(0), 1.0 / beta, s);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < m; ++j) {
      apply_M(P, o, val, vec, Vj(j), Zj(j), n, s);
      op_apply(A, o, alg, val, vec, acc, 1.0, Zj(j), 0.0, w, w, W, s);
      gs_dots(j + 1, h1);
      gs_update(V, j + 1, h1, w);
      gs_dots(j + 1, h2);
      gs_update(V, j + 1, h2, w);
      dot_dev(n, vec, w, w, part, hn, s);
      LA_HIP(hipMemcpyAsync(hbuf, h1, sizeof(double) * (2 * cap + 1), hipMemcpyDeviceToHost, s));
      LA_HIP(hipStreamSynchronize(s));
      for (int i = 0; i <= j; ++i) H[static_cast<size_t>(i) * m + j] = hbuf[i] + hbuf[cap + i];
      const double hnorm = std::sqrt(hbuf[2 * cap]);
      H[static_cast<size_t>(j + 1) * m + j] = hnorm;
      for (int i = 0; i < j; ++i) {
        const double a = H[static_cast<size_t>(i) * m + j], c = H[static_cast<size_t>(i + 1) * m + j];
        H[static_cast<size_t>(i) * m + j] = cs[i] * a + sn[i] * c;
        H[static_cast<size_t>(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
      }
      const double a = H[static_cast<size_t>(j) * m + j], c = H[static_cast<size_t>(j + 1) * m + j];
      const double rr = std::sqrt(a * a + c * c);
      cs[j] = rr == 0 ? 1.0 : a / rr;
      sn[j] = rr == 0 ? 0.0 : c / rr;
      H[static_cast<size_t>(j) * m + j] = rr;
      H[static_cast<size_t>(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      ++total;
      k = j + 1;
      const double est = std::fabs(g[j + 1]);
      R.last_res = est;
      if (o.print_level == 1) std::fprintf(stderr, "%3d = %lg/%lg=%lg\n", total, est, nb, est / nb);
      if (est / nb < rel_tol || est < abs_tol) { R.converged = true; done = true; break; }
      if (hnorm < 1.0e-16) { R.converged = true; done = true; break; }
      if (total >= maxiter) { done = true; break; }
      copy_vec(n, vec, w, vec, Vj(j + 1), 1.0 / hnorm, s);
    }
    for (int i = k - 1; i >= 0; --i) {
      double acc_y = g[i];
      for (int l = i + 1; l < k; ++l) acc_y -= H[static_cast<size_t>(i) * m + l] * y[l];
      y[i] = acc_y / H[static_cast<size_t>(i) * m + i];
    }
    // x += Z y  (multiaxpy subtracts, so pass -y)
    for (int l = 0; l < k; ++l) hbuf[l] = -y[l];
    LA_HIP(hipMemcpyAsync(h1, hbuf, sizeof(double) * k, hipMemcpyHostToDevice, s));
    gs_update(Z, k, h1, x);
    LA_HIP(hipStreamSynchronize(s));
  }
  R.iters = total;
  return R;
}

// Mixed-precision PCG with reliable updates (see the file comment). x: FP64
// device vector holding x0; vd: the inner matrix precision. Counts the true
// residual evaluations in *updates.
CgResult pcg_reliable(Mat &A, Prec *P, const la_solver_opts &o, la_dtype vd, const double *b,
                      double *x, double nb, Work &W, hipStream_t s, int *updates) {
  const int64_t n = A.rows;
  const la_spmv_alg alg = static_cast<la_spmv_alg>(o.spmv_alg);
  void *r = W.v32[0].p, *z = W.v32[1].p, *p = W.v32[2].p, *q = W.v32[3].p, *xl = W.v32[4].p;
  double *r64 = W.v64[5].as<double>();
  double *part = W.sc.partials.as<double>();
  double *p_pq = part, *p_rr = part + kRedBlocks;
  double *p_rz[2] = {part + 3 * kRedBlocks, part + 4 * kRedBlocks};
  const double goal = o.abs_tol > 0 ? std::max(o.rel_tol * nb, o.abs_tol) : o.rel_tol * nb;
  const double delta = o.reliable_delta;
  // x += nb * xl; xl = 0; r64 = b - A x; r = r64 / nb (FP32); returns ||r64||
  bool first = true;
  auto replace = [&]() {
    if (!first) add_scaled(n, LA_F32, xl, nb, x, s);
    first = false;
    fill_zero(n, LA_F32, xl, s);
    spmv(A, alg, LA_F64, LA_F64, LA_F64, -1.0, x, 1.0, b, r64, s);
    const double rn = norm2(n, LA_F64, r64, W.sc, s);
    copy_vec(n, LA_F64, r64, LA_F32, r, 1.0 / nb, s);
    ++*updates;
    if (o.print_level == 1) std::fprintf(stderr, "true residual: %lg/%lg=%lg\n", rn, nb, rn / nb);
    return rn;
  };
  CgResult R;
  double rn = replace();
  R.last_res = rn;
  if (rn < goal) {
    R.converged = true;
    return R;
  }
  int cur = 0;
  apply_M(P, o, vd, LA_F32, r, z, n, s);
  copy_vec(n, LA_F32, z, LA_F32, p, 1.0, s);
  dot_partials(n, LA_F32, r, z, p_rz[cur], s);
  double rmax = rn / nb;  // in units of ||b||, like r
  bool folded = true;     // xl holds nothing x lacks
  for (int it = 1; it <= o.max_iters; ++it) {
    spmv(A, alg, vd, LA_F32, LA_F32, 1.0, p, 0.0, q, q, s);
    dot_partials(n, LA_F32, p, q, p_pq, s);
    cg_update_xr(n, LA_F32, xl, r, p, q, p_rz[cur], p_pq, p_rr, s);
    folded = false;
    const uint64_t tk = report_sum(W.sc, p_rr, s);
    // next direction queued before the test is read (see pcg_core); after a
    // replacement it stays the one built from the recurrence residual, a
    // valid search direction, while the residual itself is the true one
    apply_M(P, o, vd, LA_F32, r, z, n, s);
    dot_partials(n, LA_F32, r, z, p_rz[cur ^ 1], s);
    cg_update_p(n, LA_F32, p, z, p_rz[cur ^ 1], p_rz[cur], s);
    const double rr = std::sqrt(wait_report(W.sc, tk, s));
    R.iters = it;
    if (o.print_level == 1) std::fprintf(stderr, "%3d = %lg (recurrence, relative)\n", it, rr);
    if (!std::isfinite(rr)) break;
    rmax = std::max(rmax, rr);
    if (rr < delta * rmax || rr * nb < goal) {
      rn = replace();
      folded = true;
      R.last_res = rn;
      if (rn < goal) {
        R.converged = true;
        break;
      }
      if (!std::isfinite(rn)) break;
      rmax = rn / nb;
    }
    cur ^= 1;
  }
  if (!folded) add_scaled(n, LA_F32, xl, nb, x, s);  // x keeps the last iterate
  LA_HIP(hipStreamSynchronize(s));
  return R;
}

using clk = std::chrono::steady_clock;

template <typename Core>
int run_solver(Mat &A, Prec *P, const la_solver_opts &o, const double *b, double *x,
               la_solver_result *res, Core core, bool is_pcg) {
  LA_CHECK_ARG(o.precision >= LA_PREC_FP64 && o.precision <= LA_PREC_MIXED_F32_BF16P,
               "unknown precision policy");
  LA_CHECK_ARG(!(o.reliable_delta >= 1.0), "reliable_delta must be below 1");
  hipStream_t s = A.stream;
  const int64_t n = A.rows;
  Work &W = work_for(n);
  la_solver_result out{};
  const double nb = norm2(n, LA_F64, b, W.sc, s);
  const auto t0 = clk::now();
  if (nb == 0.0) {  // zero right-hand side: x = 0 solves it exactly
    fill_zero(n, LA_F64, x, s);
    LA_HIP(hipStreamSynchronize(s));
    out.converged = 1;
    out.outer_iters = 1;
  } else if (o.precision == LA_PREC_FP64) {
    CgResult R = core(LA_F64, LA_F64, static_cast<const void *>(b), static_cast<void *>(x), nb,
                      o.rel_tol, o.abs_tol, o.max_iters, W);
    out.iters = R.iters;
    out.outer_iters = 1;
    out.converged = R.converged;
  } else if (is_pcg && o.reliable_delta > 0) {
    const la_dtype vd = inner_val_dtype(o.precision);
    if (!A.vals.has[vd]) A.vals.add(vd, s);
    if (P) prec_add_dtype(*P, inner_prec_dtype(o.precision));
    int updates = 0;
    CgResult R = pcg_reliable(A, P, o, vd, b, x, nb, W, s, &updates);
    out.iters = R.iters;
    out.outer_iters = updates;
    out.converged = R.converged;
  } else {
    const la_dtype vd = inner_val_dtype(o.precision);
    if (!A.vals.has[vd]) A.vals.add(vd, s);
    if (P) prec_add_dtype(*P, inner_prec_dtype(o.precision));
    double *r64 = W.v64[5].as<double>();
    float *rhs = W.v32[5].as<float>();
    // the inner solution lives in a dedicated FP32 vector outside the ones
    // the inner core uses (v32[0..4])
    DevBuf d32(sizeof(float) * std::max<int64_t>(n, 1));
    int total = 0, outer = 0;
    bool conv = false;
    for (outer = 0; outer < std::max(1, o.max_outer); ++outer) {
      op_apply(A, o, static_cast<la_spmv_alg>(o.spmv_alg), LA_F64, LA_F64, LA_F64, -1.0, x, 1.0,
               b, r64, W, s);
      const double rn = norm2(n, LA_F64, r64, W.sc, s);
      if (o.print_level == 1) std::fprintf(stderr, "outer %d: %lg/%lg=%lg\n", outer, rn, nb, rn / nb);
      if (rn / nb < o.rel_tol || rn < o.abs_tol) { conv = true; break; }
      if (total >= o.max_iters) break;
      copy_vec(n, LA_F64, r64, LA_F32, rhs, 1.0 / rn, s);
      fill_zero(n, LA_F32, d32.p, s);
      // the inner solve needs no more than the reduction that brings the
      // outer residual to the target (with a margin of 4 for the low-
      // precision operator); at most inner_tol, at least a factor 10
      const double goal = o.abs_tol > 0 ? std::max(o.rel_tol * nb, o.abs_tol) : o.rel_tol * nb;
      const double inner = std::min(0.1, std::max(o.inner_tol, 0.25 * goal / rn));
      CgResult R = core(vd, LA_F32, static_cast<const void *>(rhs), d32.p, 1.0, inner, 0.0,
                        o.max_iters - total, W);
      total += R.iters;
      add_scaled(n, LA_F32, d32.p, rn, x, s);
    }
    out.iters = total;
    out.outer_iters = outer;
    out.converged = conv;
  }
  LA_HIP(hipStreamSynchronize(s));
  out.seconds = std::chrono::duration<double>(clk::now() - t0).count();
  // final true residual in FP64
  double *t = W.v64[4].as<double>();
  op_apply(A, o, LA_SPMV_AUTO, LA_F64, LA_F64, LA_F64, -1.0, x, 1.0, b, t, W, s);
  const double tr = norm2(n, LA_F64, t, W.sc, s);
  out.rel_residual = nb > 0 ? tr / nb : 0.0;
  if (o.precision == LA_PREC_FP64 && out.converged && o.stop_rule == LA_STOP_RECURRENCE &&
      !(out.rel_residual < o.rel_tol || tr < o.abs_tol * 1.0000001))
    out.converged = 0;  // recurrence said yes, the true residual says no
  for (Prec *Q : {P, tl_right}) {
    if (Q && (trsv_check_timeout(*Q->lower, s) || trsv_check_timeout(*Q->upper, s)))
      throw la_error(LA_ERR_TIMEOUT, "sync-free triangular solve exceeded its spin bound");
  }
  if (res) *res = out;
  return out.converged ? LA_OK : LA_ERR_NOT_CONVERGED;
}
}  // namespace

// A^T x = b with the same loops: the product uses the transposed matrix
// (built once, stable transpose) and the preconditioner its transposed apply.
struct TransposeScope {
  explicit TransposeScope(bool on) { tl_transpose = on; }
  ~TransposeScope() { tl_transpose = false; }
};
Mat &transposed_of(Mat &A);

int pcg_solve(Mat &A, Prec *P, const la_solver_opts &o, const double *b, double *x,
              la_solver_result *res) {
  LA_CHECK_ARG(A.rows == A.cols, "PCG needs a square matrix");
  LA_CHECK_ARG(!P || P->n == A.rows, "preconditioner size mismatch");
  hipStream_t s = A.stream;
  return run_solver(A, P, o, b, x, res,
                    [&](la_dtype val, la_dtype vec, const void *bb, void *xx, double nb,
                        double rt, double at, int mi, Work &W) {
                      const bool true_res = vec == LA_F64 && o.stop_rule == LA_STOP_TRUE_RESIDUAL;
                      return pcg_core(A, P, o, val, vec, bb, xx, nb, rt, at, mi, true_res, W, s);
                    },
                    true);
}

int fgmres_solve(Mat &A, Prec *P, const la_solver_opts &o, const double *b, double *x,
                 la_solver_result *res) {
  LA_CHECK_ARG(A.rows == A.cols, "FGMRES needs a square matrix");
  LA_CHECK_ARG(!P || P->n == A.rows, "preconditioner size mismatch");
  hipStream_t s = A.stream;
  // restart 0: unrestarted, up to max_iters (mkl_solver.cpp:267-269, 293-294)
  LA_CHECK_ARG(o.restart >= 0 && o.max_iters >= 0, "negative restart or max_iters");
  const int m = std::max(1, o.restart > 0 ? std::min(o.restart, std::max(o.max_iters, 1))
                                          : o.max_iters);
  return run_solver(A, P, o, b, x, res,
                    [&](la_dtype val, la_dtype vec, const void *bb, void *xx, double nb,
                        double rt, double at, int mi, Work &W) {
                      return fgmres_core(A, P, o, val, vec, bb, xx, nb, rt, at, mi, m, W, s);
                    },
                    false);
}

// FGMRES on A R^-1 u = b (mkl_fgmres_solver with a right operator R,
// mkl_solver.cpp:319-324): returns u, as the reference does; the solution
// of A x = b is R^-1 u. FP64 only.
int fgmres_solve_right(Mat &A, Prec *P, Prec *R, const la_solver_opts &o, const double *b,
                       double *x, la_solver_result *res) {
  LA_CHECK_ARG(R && R->n == A.rows, "right operator size mismatch");
  LA_CHECK_ARG(o.precision == LA_PREC_FP64, "a right operator runs in FP64 only");
  struct RightScope {
    explicit RightScope(Prec *r) { tl_right = r; }
    ~RightScope() { tl_right = nullptr; }
  } scope(R);
  return fgmres_solve(A, P, o, b, x, res);
}

}  // namespace laamd

namespace laamd {
Mat &transposed_of(Mat &A) { return ensure_transposed(A, A.stream); }

// CG presumes a symmetric A, for which A^T = A: on a structurally symmetric
// pattern the transposed solve runs on A itself (no transposed copy); the
// preconditioner's transposed apply is still used (an IC apply is its own
// transpose).
int pcg_solve_transpose(Mat &A, Prec *P, const la_solver_opts &o, const double *b, double *x,
                        la_solver_result *res) {
  Mat &At = mat_symmetric_pattern(A) ? A : transposed_of(A);
  TransposeScope scope(true);
  return pcg_solve(At, P, o, b, x, res);
}

int fgmres_solve_transpose(Mat &A, Prec *P, const la_solver_opts &o, const double *b, double *x,
                           la_solver_result *res) {
  Mat &At = transposed_of(A);
  TransposeScope scope(true);
  return fgmres_solve(At, P, o, b, x, res);
}
}  // namespace laamd
