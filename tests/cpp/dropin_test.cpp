// Drop-in C++ layer test, written the way the reference's gtest files use
// the same interfaces (tests/spmv_test.cpp, tests/triangular_solve_test.cpp,
// execs/linearSolve.cpp): SpMV functors, SPMV<>, OptimizedTriangularSolve,
// IC(0)/ILU(k) preconditioners and the PCG/FGMRES solvers on a 2-D Poisson
// matrix, against plain serial loops. Prints PASS/FAIL per check; exit code
// is the number of failures.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "laamd/dropin.hpp"

using namespace laamd;
static int failures = 0;
static void expect(bool ok, const char *name, double v) {
  std::printf("%s %s (%.3e)\n", ok ? "PASS" : "FAIL", name, v);
  if (!ok) ++failures;
}

static double rel(const std::vector<double> &a, const std::vector<double> &b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return std::sqrt(num / (den > 0 ? den : 1));
}

int main() {
  const int nx = 48, n = nx * nx;
  std::vector<int> ai(n + 1);
  int64_t nnz = 0;
  detail::check(la_gen_poisson2d(nx, nx, ai.data(), nullptr, nullptr, &nnz));
  std::vector<int> aj(nnz);
  std::vector<double> av(nnz);
  detail::check(la_gen_poisson2d(nx, nx, ai.data(), aj.data(), av.data(), &nnz));
  std::vector<double> b(n), ref(n), x(n);
  for (int i = 0; i < n; ++i) b[i] = std::sin(0.1 * i);
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (int k = ai[i]; k < ai[i + 1]; ++k) s += av[k] * b[aj[k]];
    ref[i] = s;
  }
  // SpMV functors (spmv.hpp)
  matrix_utils::SerialSPMV serial;
  serial(n, 0, ai.data(), aj.data(), av.data(), b.data(), x.data(), 1., 0.);
  expect(rel(x, ref) < 1e-14, "SerialSPMV", rel(x, ref));
  matrix_utils::ALBUSSPMV<> albus(4);
  albus.preprocess(n, 0, ai.data(), aj.data(), av.data());
  albus(n, 0, ai.data(), aj.data(), av.data(), b.data(), x.data(), 1., 0.);
  expect(rel(x, ref) < 1e-14, "ALBUSSPMV", rel(x, ref));
  matrix_utils::CSRMatrixVec<int, int, double> M;
  M.rows = M.cols = n;
  M.ai = ai;
  M.aj = aj;
  M.av = av;
  matrix_utils::SPMV<decltype(M), matrix_utils::SegSumSPMV<>> spmv;
  spmv.setMatrix(&M);
  spmv.preprocess();
  std::vector<double> y(n, 1.0);
  spmv(b.data(), y.data(), 2.0, -1.0);
  std::vector<double> ref2(n);
  for (int i = 0; i < n; ++i) ref2[i] = 2.0 * ref[i] - 1.0;
  expect(rel(y, ref2) < 1e-14, "SPMV<SegSum> alpha/beta", rel(y, ref2));

  // values changed in place between calls (a refactor / Newton loop): every
  // form reads the new av, as the reference's operator() does
  {
    std::vector<double> av2 = av;
    for (size_t k = 0; k < av2.size(); ++k) av2[k] *= 1.0 + 0.25 * std::cos(0.3 * k);
    std::vector<double> ref3(n);
    for (int i = 0; i < n; ++i) {
      double s = 0;
      for (int k = ai[i]; k < ai[i + 1]; ++k) s += av2[k] * b[aj[k]];
      ref3[i] = s;
    }
    std::vector<double> avm = av;  // the array the functors were given
    matrix_utils::ALBUSSPMV<> a2(1);
    a2.preprocess(n, 0, ai.data(), aj.data(), avm.data());
    a2(n, 0, ai.data(), aj.data(), avm.data(), b.data(), x.data(), 1., 0.);
    expect(rel(x, ref) < 1e-14, "ALBUSSPMV before the value change", rel(x, ref));
    avm = av2;  // same pointer, new values
    a2(n, 0, ai.data(), aj.data(), avm.data(), b.data(), x.data(), 1., 0.);
    expect(rel(x, ref3) < 1e-14, "ALBUSSPMV reads values changed in place", rel(x, ref3));
    matrix_utils::CSRMatrixVec<int, int, double> M2 = M;
    matrix_utils::SPMV<decltype(M2), matrix_utils::ALBUSSPMV<>> s2;
    s2.setMatrix(&M2);
    s2.preprocess();
    s2(b.data(), x.data());
    expect(rel(x, ref) < 1e-14, "SPMV<ALBUS> before the value change", rel(x, ref));
    M2.av = av2;
    s2(b.data(), x.data());
    expect(rel(x, ref3) < 1e-14, "SPMV<ALBUS> reads values changed in place", rel(x, ref3));
    detail::DeviceBuffer db(sizeof(double) * n), dx(sizeof(double) * n);
    db.upload(b.data(), sizeof(double) * n);
    M2.av = av;
    s2.apply_device(static_cast<const double *>(db.get()), static_cast<double *>(dx.get()));
    dx.download(x.data(), sizeof(double) * n);
    expect(rel(x, ref) < 1e-14, "SPMV<ALBUS>::apply_device reads the values again", rel(x, ref));
  }

  // triangular solves (triangle_solve.hpp) on L/D/U of A
  matrix_utils::CSRMatrixVec<int, int, double> Lm, Um;
  std::vector<double> D;
  matrix_utils::SplitLDU(n, 0, ai.data(), aj.data(), av.data(), Lm, D, Um);
  std::vector<double> xs(n), xl(n);
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (int k = Lm.ai[i]; k < Lm.ai[i + 1]; ++k) s += Lm.av[k] * xs[Lm.aj[k]];
    xs[i] = b[i] - s;
  }
  matrix_utils::OptimizedTriangularSolve<matrix_utils::FBSubstitutionType::Barrier,
                                         matrix_utils::L>
      fw;
  fw.analysis(n, 0, Lm.AI(), Lm.AJ(), Lm.AV());
  for (int rep = 0; rep < 3; ++rep) fw(b.data(), xl.data());
  expect(rel(xl, xs) < 1e-14, "OptimizedTriangularSolve<Barrier,L>", rel(xl, xs));
  expect(fw.levels() == 2 * nx - 1, "levels == 2nx-1", fw.levels());
  matrix_utils::OptimizedTriangularSolve<matrix_utils::FBSubstitutionType::NoBarrier,
                                         matrix_utils::U>
      bw;
  bw.analysis(n, 0, Um.AI(), Um.AJ(), Um.AV(), D.data());
  for (int i = n - 1; i >= 0; --i) {
    double s = 0;
    for (int k = Um.ai[i]; k < Um.ai[i + 1]; ++k) s += Um.av[k] * xs[Um.aj[k]];
    xs[i] = (b[i] - s) / D[i];
  }
  bw(b.data(), xl.data());
  expect(rel(xl, xs) < 1e-14, "OptimizedTriangularSolve<NoBarrier,U>", rel(xl, xs));

  // IC(0)-PCG as execs/linearSolve.cpp builds it
  mkl_wrapper::mkl_sparse_mat A(n, n, ai, aj, av);
  mkl_wrapper::mkl_sparse_mat_sym Asym(A);
  auto prec = std::make_shared<mkl_wrapper::incomplete_cholesky_k>();
  prec->set_level(0);
  prec->shift(true);
  bool ok = prec->symbolic_factorize(&Asym) && prec->numeric_factorize(&Asym);
  expect(ok, "incomplete_cholesky_k factorize", 0);
  expect(prec->mkl_base() == 1 && prec->get_ai()[0] == 1, "IC factor left 1-based",
         prec->get_ai()[0]);
  mkl_wrapper::mkl_pcg_solver pcg(&Asym, prec);
  pcg.set_rel_tol(1e-10);
  pcg.set_max_iters(1000);
  std::vector<double> sol(n, 0.0), rhs(ref);
  ok = pcg.solve(rhs.data(), sol.data());
  expect(ok && rel(sol, b) < 1e-7, "mkl_pcg_solver + IC(0)", rel(sol, b));
  expect(pcg.last_result().rel_residual < 1e-10, "PCG true residual", pcg.last_result().rel_residual);
  // extensions with no counterpart in the reference, on the same solver:
  // Jacobi-sweep triangular solves and mixed BF16 precision with reliable
  // updates, still to an FP64 true residual of 1e-10
  prec->set_trsv_mode(LA_TRSV_JACOBI);
  prec->set_sweeps(3);
  pcg.set_precision(LA_PREC_MIXED_BF16);
  pcg.set_reliable_delta(0.1);
  std::fill(sol.begin(), sol.end(), 0.0);
  ok = pcg.solve(rhs.data(), sol.data());
  expect(ok && pcg.last_result().rel_residual < 1e-10, "PCG mixed BF16 + Jacobi sweeps",
         pcg.last_result().rel_residual);
  expect(pcg.last_result().outer_iters > 1, "reliable updates replaced the residual",
         pcg.last_result().outer_iters);
  pcg.set_precision(LA_PREC_FP64);
  prec->set_trsv_mode(LA_TRSV_SYNCFREE);

  // incomplete_cholesky_k_2 factors the same IC(0) pattern with the rows of
  // each column visited in the opposite order: the factors agree to rounding
  {
    auto p2 = std::make_shared<mkl_wrapper::incomplete_cholesky_k_2>();
    p2->set_level(0);
    ok = p2->symbolic_factorize(&Asym) && p2->numeric_factorize(&Asym);
    double dmax = ok && p2->nnz() == prec->nnz() ? 0.0 : 1.0;
    for (int q = 0; ok && q < prec->nnz(); ++q) {
      const double a = prec->get_av()[q], c = p2->get_av()[q];
      dmax = std::max(dmax, std::fabs(a - c) / std::fabs(c));
    }
    expect(dmax < 1e-13, "incomplete_cholesky_k_2 vs incomplete_cholesky_k", dmax);
  }
  // a second numeric phase on the 1-based factor (new values, same pattern)
  // gives the same factor; to_zero_based shifts the indices back
  {
    std::vector<double> v0(prec->get_av().get(), prec->get_av().get() + prec->nnz());
    ok = prec->numeric_factorize(&Asym);
    bool same = ok && prec->get_ai()[0] == 1;
    for (int q = 0; same && q < prec->nnz(); ++q) same = prec->get_av()[q] == v0[q];
    expect(same, "numeric_factorize again on the 1-based factor", 0);
    prec->to_zero_based();
    expect(prec->mkl_base() == 0 && prec->get_ai()[0] == 0 && prec->get_aj()[0] == 0,
           "to_zero_based", prec->get_ai()[0]);
  }
  // additive shift (incomplete_cholesky.cpp:207-266): [[-1, .5], [.5, 2]]
  // starts at s = 1e-3 + 1, fails on the second pivot, retries at 2.002;
  // r_11 = sqrt(-1 + 2.002)
  {
    mkl_wrapper::mkl_sparse_mat S2(2, 2, std::vector<int>{0, 2, 4}, std::vector<int>{0, 1, 0, 1},
                                   std::vector<double>{-1.0, 0.5, 0.5, 2.0});
    mkl_wrapper::mkl_sparse_mat_sym S2sym(S2);
    mkl_wrapper::incomplete_cholesky_k p3;
    p3.set_level(0);
    p3.shift(true);
    ok = p3.symbolic_factorize(&S2sym) && p3.numeric_factorize(&S2sym);
    const double want = std::sqrt(-1.0 + 2 * (1e-3 + 1.0));
    expect(ok && p3.get_av()[0] == want, "incomplete_cholesky_k additive shift retry",
           ok ? p3.get_av()[0] - want : 1.0);
    mkl_wrapper::incomplete_cholesky_k p4;
    p4.set_level(0);
    ok = p4.symbolic_factorize(&S2sym) && p4.numeric_factorize(&S2sym);
    expect(!ok, "incomplete_cholesky_k without shift rejects a pivot <= 0", 0);
  }

  // ILU(1)-FGMRES through the factory, the "gmres+iluk" registration of
  // execs/test1.cpp:61-77
  auto &factory = utils::singleton<mkl_wrapper::solver_factory>::instance();
  std::shared_ptr<mkl_wrapper::incomplete_lu_k> ilu;
  factory.reg("gmres+iluk", [&](mkl_wrapper::mkl_sparse_mat &m) {
    ilu = std::make_shared<mkl_wrapper::incomplete_lu_k>();
    ilu->set_level(1);
    ilu->symbolic_factorize(&m);
    ilu->numeric_factorize(&m);
    auto s = std::make_unique<mkl_wrapper::mkl_fgmres_solver>(&m, ilu);
    s->set_rel_tol(1e-10);
    s->set_restart_steps(30);
    return mkl_wrapper::solver_factory::solver_ptr(std::move(s));
  });
  auto gm = factory.create("gmres+iluk", A);
  std::fill(sol.begin(), sol.end(), 0.0);
  ok = gm && gm->solve(rhs.data(), sol.data());
  expect(ok && rel(sol, b) < 1e-7, "factory gmres+iluk", rel(sol, b));
  expect(factory.create("no-such-solver", A) == nullptr, "factory unknown name", 0);

  // FGMRES with a right operator (mkl_solver.cpp:319-324): returns u with
  // A R^-1 u = b; x = R^-1 u solves A x = b
  {
    auto R = std::make_shared<mkl_wrapper::incomplete_lu_k>();
    R->set_level(0);
    R->symbolic_factorize(&A);
    R->numeric_factorize(&A);
    mkl_wrapper::mkl_fgmres_solver fr(&A, nullptr, R);
    fr.set_rel_tol(1e-11);
    fr.set_restart_steps(30);
    fr.set_max_iters(500);
    std::vector<double> u(n, 0.0), xr(n, 0.0);
    ok = fr.solve(rhs.data(), u.data());
    R->solve(u.data(), xr.data());
    expect(ok && rel(xr, b) < 1e-7, "mkl_fgmres_solver with right operator", rel(xr, b));
  }

  // symmetric scaling (mkl_sparse_mat.cpp:517-544): D A D with D from the
  // row norms has unit-ish rows; the product follows the scaled values
  {
    mkl_wrapper::mkl_sparse_mat S(n, n, ai, aj, av);
    auto d = S.rowwiseSqrtNorm();
    S.DtAD(d);
    std::vector<double> y(n), yref(n);
    S.mult_vec(b.data(), y.data());
    for (int i = 0; i < n; ++i) {
      double acc = 0;
      for (int q = ai[i]; q < ai[i + 1]; ++q) acc += av[q] * d[i] * d[aj[q]] * b[aj[q]];
      yref[i] = acc;
    }
    expect(rel(y, yref) < 1e-13, "rowwiseSqrtNorm + DtAD product", rel(y, yref));
  }
  std::printf("%d failure(s)\n", failures);
  return failures;
}
