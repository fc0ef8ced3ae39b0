// ILU(k) / IC(k) numeric factorisation on the device.
//
// The symbolic phase stays on the host (iluk_symbolic / icc_symbolic, bit-
// exact with the reference); the numeric phase runs on gfx950 with the
// reference's arithmetic in the reference's order, so the factor is bitwise
// the host's:
//  * ILU (mkl_wrapper/incomplete_lu.cpp:277-325, row IKJ): for each row i,
//    for each k < i in the row's pattern in ascending order,
//    l_ik = a_ik / u_kk, then a_ij -= l_ik * u_kj for every j > k of row k's
//    upper part that is in row i's pattern.
//  * IC (sparse_mat_op/precond.hpp:615-693, up-looking on the upper
//    pattern): the diagonal is scaled by (1 + omega); for each row k with an
//    entry in column j, in the reference's linked-list order (recorded on
//    the host once per pattern), u_j* -= u_kj * u_k* over the merged
//    columns; then u_jj = sqrt(pivot) and the row is divided by it; a
//    pivot < 0 fails.
//  * IC with the additive shift (mkl_wrapper/incomplete_cholesky.cpp:
//    195-274, incomplete_cholesky_k): the same updates with the shift added
//    to the diagonal, the rows k of each column taken first-in first-out
//    (that class's column vectors), and a pivot <= 0 failing.
// Products and differences are rounded separately (no fused multiply-add),
// as in the reference's x86-64 build.
//
// Rows are processed level by level (one launch per level of the
// factorisation's dependency DAG, which is the lower triangular solve's),
// one 64-lane workgroup per row: the row's columns and values sit in LDS;
// for each eliminating row k the lanes take k's entries, find their column
// in the row by binary search and update it. A row longer than kRowCap is
// done by one lane in global memory, in the host's loop order.
#include <algorithm>
#include <climits>
#include <cmath>

#include "dev_util.hpp"
#include "device.hpp"
#include "host_algos.hpp"

// every product and difference in this file rounds on its own (the
// library's rounding intrinsics do not stop the fusion once inlined; the
// pragma on plain operators does)
#pragma clang fp contract(off)

namespace laamd {

struct NumericPlan {
  bool ic = false;
  int n = 0;
  int64_t nnzA = 0, nnzF = 0;
  double omega = 0.0;
  bool additive = false;   // IC: incomplete_cholesky_k's arithmetic, omega = the shift
  DevBuf fai, faj, fu;     // pattern (0-based); ILU: first upper position per row
  DevBuf a2f;              // per entry of A: its position in the factor, or -1
  DevBuf fav, diag;        // factor values; ILU: u_ii per row
  DevBuf rows;             // rows in dependency-level order
  std::vector<int> level_ptr;
  DevBuf pptr, pk, pq;     // IC: per row j the (k, q) eliminations, host order
  DevBuf err;              // row + 1 of a failed pivot, else 0
  std::vector<int> h_fai, h_faj;
};
void NumericPlanDeleter::operator()(NumericPlan *p) const { delete p; }

namespace {
constexpr int kRowCap = 512;

__global__ void k_scatter(int64_t m, const int64_t *__restrict__ a2f,
                          const double *__restrict__ av, double *fav) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < m) {
    const int64_t p = a2f[t];
    if (p >= 0) fav[p] = av[t];
  }
}

__global__ void k_scale_first(int n, const int *__restrict__ fai, double *fav, double f) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n && fai[j + 1] > fai[j]) fav[fai[j]] = fav[fai[j]] * f;
}
__global__ void k_shift_first(int n, const int *__restrict__ fai, double *fav, double a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n && fai[j + 1] > fai[j]) fav[fai[j]] = fav[fai[j]] + a;
}

__global__ void k_fill_zero(int64_t m, double *v) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < m;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    v[t] = 0.0;
}

__device__ __forceinline__ int find_col(const int *cols, int lo, int hi, int c) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cols[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ void flag_row(int *err, int row) {
  atomicCAS(err, 0, row + 1);
}

// one row of ILU by one lane, in global memory (rows longer than kRowCap)
__device__ void ilu_row_serial(int i, const int *fai, const int *faj, const int *fu, double *fav,
                               double *diag, int *err) {
  const int fb = fai[i], fe = fai[i + 1];
  int t = fb;
  for (; t < fe && faj[t] != i; ++t) {
    const int k = faj[t];
    fav[t] = fav[t] / diag[k];
    const double lik = fav[t];
    int p = t, q = fu[k];
    const int qe = fai[k + 1];
    while (q != qe && p != fe) {
      if (faj[p] == faj[q]) {
        fav[p] = fav[p] - lik * fav[q];
        ++p;
        ++q;
      } else if (faj[p] < faj[q]) {
        ++p;
      } else {
        ++q;
      }
    }
  }
  if (t == fe) flag_row(err, i);
  else diag[i] = fav[t];
}

__global__ __launch_bounds__(64) void k_ilu_rows(int k0, const int *__restrict__ rows,
                                                 const int *__restrict__ fai,
                                                 const int *__restrict__ faj,
                                                 const int *__restrict__ fu, double *fav,
                                                 double *diag, int *err) {
  __shared__ int cols[kRowCap];
  __shared__ double vals[kRowCap];
  const int i = rows[k0 + blockIdx.x];
  const int lane = threadIdx.x;
  const int fb = fai[i], len = fai[i + 1] - fb;
  if (len > kRowCap) {
    if (lane == 0) ilu_row_serial(i, fai, faj, fu, fav, diag, err);
    return;
  }
  for (int t = lane; t < len; t += 64) {
    cols[t] = faj[fb + t];
    vals[t] = fav[fb + t];
  }
  __syncthreads();
  int t = 0;
  for (; t < len && cols[t] != i; ++t) {
    const int k = cols[t];
    if (lane == 0) vals[t] = vals[t] / diag[k];
    __syncthreads();
    const double lik = vals[t];
    const int qe = fai[k + 1];
    for (int q = fu[k] + lane; q < qe; q += 64) {
      const int c = faj[q];
      const int p = find_col(cols, t + 1, len, c);
      if (p < len && cols[p] == c) vals[p] = vals[p] - lik * fav[q];
    }
    __syncthreads();
  }
  if (lane == 0) {
    if (t == len) flag_row(err, i);
    else diag[i] = vals[t];
  }
  for (int u = lane; u < len; u += 64) fav[fb + u] = vals[u];
}

// one row of IC by one lane, in global memory
// (strict: a zero pivot fails too)
__device__ void ic_row_serial(int j, const int *fai, const int *faj, const int *pptr,
                              const int *pk, const int *pq, double *fav, int *err, int strict) {
  const int ps = fai[j], pe = fai[j + 1];
  for (int e = pptr[j]; e < pptr[j + 1]; ++e) {
    const int k = pk[e];
    int q = pq[e];
    const int qe = fai[k + 1];
    const double ukj = fav[q];
    int t = ps;
    while (q < qe && t < pe) {
      if (faj[q] == faj[t]) {
        fav[t] = fav[t] - ukj * fav[q];
        ++t;
        ++q;
      } else if (faj[q] < faj[t]) {
        ++q;
      } else {
        ++t;
      }
    }
  }
  if (strict ? fav[ps] <= 0 : fav[ps] < 0) {
    flag_row(err, j);
    return;
  }
  const double d = sqrt(fav[ps]);
  fav[ps] = d;
  for (int t = ps + 1; t < pe; ++t) fav[t] = fav[t] / d;
}

__global__ __launch_bounds__(64) void k_ic_rows(int k0, const int *__restrict__ rows,
                                                const int *__restrict__ fai,
                                                const int *__restrict__ faj,
                                                const int *__restrict__ pptr,
                                                const int *__restrict__ pk,
                                                const int *__restrict__ pq, double *fav,
                                                int *err, int strict) {
  __shared__ int cols[kRowCap];
  __shared__ double vals[kRowCap];
  const int j = rows[k0 + blockIdx.x];
  const int lane = threadIdx.x;
  const int ps = fai[j], len = fai[j + 1] - ps;
  if (len > kRowCap) {
    if (lane == 0) ic_row_serial(j, fai, faj, pptr, pk, pq, fav, err, strict);
    return;
  }
  for (int t = lane; t < len; t += 64) {
    cols[t] = faj[ps + t];
    vals[t] = fav[ps + t];
  }
  __syncthreads();
  for (int e = pptr[j]; e < pptr[j + 1]; ++e) {
    const int k = pk[e], q0 = pq[e], qe = fai[k + 1];
    const double ukj = fav[q0];
    for (int q = q0 + lane; q < qe; q += 64) {
      const int c = faj[q];
      const int t = find_col(cols, 0, len, c);
      if (t < len && cols[t] == c) vals[t] = vals[t] - ukj * fav[q];
    }
    __syncthreads();
  }
  const double piv = vals[0];
  if (strict ? piv <= 0 : piv < 0) {
    if (lane == 0) flag_row(err, j);
    return;
  }
  const double d = sqrt(piv);
  for (int u = lane; u < len; u += 64) fav[ps + u] = u == 0 ? d : vals[u] / d;
}

template <typename T>
void upload_vec(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
  b.alloc(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (!v.empty()) b.upload(v.data(), sizeof(T) * v.size(), s);
}

// strict lower (cols < row) or upper (cols > row) part of a 0-based pattern,
// with each entry's position in the pattern
void strict_part(int n, const std::vector<int> &pi, const std::vector<int> &pj, bool upper,
                 std::vector<int> &ti, std::vector<int> &tj, std::vector<int64_t> &src) {
  ti.assign(static_cast<size_t>(n) + 1, 0);
  tj.clear();
  src.clear();
  for (int i = 0; i < n; ++i) {
    for (int p = pi[i]; p < pi[i + 1]; ++p) {
      const int c = pj[p];
      if (upper ? c > i : c < i) {
        tj.push_back(c);
        src.push_back(p);
      }
    }
    ti[i + 1] = static_cast<int>(tj.size());
  }
}

// per entry of A (0-based CSR): position of its column in the factor row,
// or -1 when the entry is not part of the factor (IC: below the diagonal)
std::vector<int64_t> map_a_to_f(int n, const int *ai, const int *aj, int base,
                                const std::vector<int> &fi, const std::vector<int> &fj,
                                bool upper_only) {
  const int64_t nnz = static_cast<int64_t>(ai[n]) - base;
  std::vector<int64_t> m(static_cast<size_t>(nnz), -1);
  for (int i = 0; i < n; ++i) {
    int p = fi[i];
    for (int64_t a = ai[i] - base; a < ai[i + 1] - base; ++a) {
      const int c = aj[a] - base;
      if (upper_only && c < i) continue;
      while (p < fi[i + 1] && fj[p] < c) ++p;
      LA_CHECK_ARG(p < fi[i + 1] && fj[p] == c, "matrix entry outside the factor's pattern");
      m[a] = p;
    }
  }
  return m;
}

void finish_plans(Prec &P, NumericPlan &N, const std::vector<int> &Li, const std::vector<int> &Lj,
                  const std::vector<int64_t> &Ls, const std::vector<int64_t> *Ld,
                  const std::vector<int> &Ui, const std::vector<int> &Uj,
                  const std::vector<int64_t> &Us, const std::vector<int64_t> &Ud, hipStream_t s) {
  const int n = N.n;
  std::vector<double> zl(std::max<size_t>(Lj.size(), 1), 0.0), zu(std::max<size_t>(Uj.size(), 1), 0.0),
      ones(static_cast<size_t>(std::max(n, 1)), 1.0);
  P.lower = make_trsv(n, 0, Li.data(), Lj.data(), zl.data(), Ld ? ones.data() : nullptr, false, s);
  P.upper = make_trsv(n, 0, Ui.data(), Uj.data(), zu.data(), ones.data(), true, s);
  trsv_set_sources(*P.lower, Ls.data(), Ld ? Ld->data() : nullptr, s);
  trsv_set_sources(*P.upper, Us.data(), Ud.data(), s);
  // the factorisation's dependency levels are the lower solve's
  upload_vec(N.rows, P.lower->h_perm, s);
  N.level_ptr = P.lower->level_ptr;
  N.fav.alloc(sizeof(double) * static_cast<size_t>(std::max<int64_t>(N.nnzF, 1)));
  N.err.alloc(sizeof(int) * 4);
  P.tmp[0].alloc(sizeof(double) * std::max(n, 1));
  P.tmp[1].alloc(sizeof(float) * std::max(n, 1));
  LA_HIP(hipStreamSynchronize(s));
}
}  // namespace

std::unique_ptr<Prec> make_prec_ilu_gpu(int n, int base, const int *ai, const int *aj, int level,
                                        hipStream_t s) {
  LA_CHECK_ARG(n >= 0 && level >= 0, "bad size or level");
  auto P = std::make_unique<Prec>();
  P->n = n;
  P->stream = s;
  P->num.reset(new NumericPlan());
  NumericPlan &N = *P->num;
  N.n = n;
  std::vector<int> fai, faj, fu;
  iluk_symbolic(n, base, ai, aj, level, fai, faj, fu);
  for (auto &v : fai) v -= base;
  for (auto &v : faj) v -= base;
  N.nnzF = fai[n];
  N.nnzA = static_cast<int64_t>(ai[n]) - base;
  // diagonal positions; ILU needs every one
  std::vector<int64_t> dpos(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    int p = fai[i];
    while (p < fai[i + 1] && faj[p] != i) ++p;
    if (p == fai[i + 1]) throw la_error(LA_ERR_NUMERIC, "ILU pattern has a row without a diagonal");
    dpos[i] = p;
  }
  std::vector<int64_t> a2f = map_a_to_f(n, ai, aj, base, fai, faj, false);
  std::vector<int> Li, Lj, Ui, Uj;
  std::vector<int64_t> Ls, Us;
  strict_part(n, fai, faj, false, Li, Lj, Ls);
  strict_part(n, fai, faj, true, Ui, Uj, Us);
  upload_vec(N.fai, fai, s);
  upload_vec(N.faj, faj, s);
  upload_vec(N.fu, fu, s);
  upload_vec(N.a2f, a2f, s);
  N.diag.alloc(sizeof(double) * std::max(n, 1));
  finish_plans(*P, N, Li, Lj, Ls, nullptr, Ui, Uj, Us, dpos, s);
  N.h_fai = std::move(fai);
  N.h_faj = std::move(faj);
  return P;
}

std::unique_ptr<Prec> make_prec_ic_gpu(int n, int base, const int *ai, const int *aj, int level,
                                       double omega, hipStream_t s, bool additive) {
  LA_CHECK_ARG(n >= 0 && level >= 0, "bad size or level");
  auto P = std::make_unique<Prec>();
  P->ic = true;
  P->n = n;
  P->stream = s;
  P->num.reset(new NumericPlan());
  NumericPlan &N = *P->num;
  N.ic = true;
  N.n = n;
  N.omega = omega;
  N.additive = additive;
  N.nnzA = static_cast<int64_t>(ai[n]) - base;
  // upper part of A with the diagonal first in each row
  std::vector<int> ui(static_cast<size_t>(n) + 1, 0), uj, dpos(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    bool has_diag = false;
    for (int64_t a = ai[i] - base; a < ai[i + 1] - base; ++a) {
      const int c = aj[a] - base;
      if (c < i) continue;
      if (c == i) {
        has_diag = true;
        dpos[i] = static_cast<int>(uj.size());
      }
      uj.push_back(c);
    }
    if (!has_diag || uj[dpos[i]] != i || dpos[i] != ui[i])
      throw la_error(LA_ERR_NUMERIC, "IC needs every diagonal entry (sorted rows)");
    ui[i + 1] = static_cast<int>(uj.size());
  }
  std::vector<int> fai, faj;
  icc_symbolic(n, 0, ui.data(), uj.data(), dpos.data(), level, fai, faj);
  N.nnzF = fai[n];
  std::vector<int64_t> a2f = map_a_to_f(n, ai, aj, base, fai, faj, true);
  // the reference's elimination order per row: icc_numeric's list walk
  // (last-in first-out) or ick_numeric's column queues (first-in first-out)
  std::vector<int> pptr(static_cast<size_t>(n) + 1, 0), pk, pq;
  if (!additive) {
    const int NONE = INT_MAX;
    std::vector<int> llist(static_cast<size_t>(n), NONE);
    std::vector<int64_t> jk(static_cast<size_t>(n), 0);
    for (int j = 0; j < n; ++j) {
      const int64_t ps = fai[j], pe = fai[j + 1];
      for (int k = llist[j]; k < j;) {
        const int64_t q = jk[k]++;
        const int64_t qe = fai[k + 1];
        const int knext = llist[k];
        if (q + 1 < qe) {
          const int c = faj[q + 1];
          llist[k] = llist[c];
          llist[c] = k;
        }
        pk.push_back(k);
        pq.push_back(static_cast<int>(q));
        k = knext;
      }
      if (pe - ps > 1) {
        const int c = faj[ps + 1];
        llist[j] = llist[c];
        llist[c] = j;
        jk[j] = ps + 1;
      }
      pptr[j + 1] = static_cast<int>(pk.size());
    }
  } else {
    const int NONE = -1;
    std::vector<int> head(static_cast<size_t>(n), NONE), tail(static_cast<size_t>(n), NONE),
        next(static_cast<size_t>(n), NONE);
    std::vector<int64_t> pos(static_cast<size_t>(n), 0);
    auto enqueue = [&](int k, int c) {
      next[k] = NONE;
      if (head[c] == NONE) head[c] = k;
      else next[tail[c]] = k;
      tail[c] = k;
    };
    for (int j = 0; j < n; ++j) {
      const int64_t ps = fai[j], pe = fai[j + 1];
      for (int k = head[j]; k != NONE;) {
        const int knext = next[k];
        const int64_t q = pos[k];
        if (q + 1 < fai[k + 1]) {
          pos[k] = q + 1;
          enqueue(k, faj[q + 1]);
        }
        pk.push_back(k);
        pq.push_back(static_cast<int>(q));
        k = knext;
      }
      head[j] = NONE;
      if (pe - ps > 1) {
        pos[j] = ps + 1;
        enqueue(j, faj[ps + 1]);
      }
      pptr[j + 1] = static_cast<int>(pk.size());
    }
  }
  // solves: upper = strict upper of R with its diagonal, lower = its transpose
  std::vector<int> Ui, Uj;
  std::vector<int64_t> Us, dsrc(static_cast<size_t>(n));
  strict_part(n, fai, faj, true, Ui, Uj, Us);
  for (int j = 0; j < n; ++j) dsrc[j] = fai[j];
  std::vector<int> Li(static_cast<size_t>(n) + 1), Lj(Uj.size());
  std::vector<double> uv(Us.begin(), Us.end()), lv(Uj.size());
  transpose(n, n, 0, Ui.data(), Uj.data(), uv.data(), Li.data(), Lj.data(), lv.data());
  std::vector<int64_t> Ls(lv.begin(), lv.end());
  upload_vec(N.fai, fai, s);
  upload_vec(N.faj, faj, s);
  upload_vec(N.a2f, a2f, s);
  upload_vec(N.pptr, pptr, s);
  upload_vec(N.pk, pk, s);
  upload_vec(N.pq, pq, s);
  finish_plans(*P, N, Li, Lj, Ls, &dsrc, Ui, Uj, Us, dsrc, s);
  N.h_fai = std::move(fai);
  N.h_faj = std::move(faj);
  return P;
}

void prec_refactor(Prec &P, const double *a_vals, hipStream_t s) {
  if (!P.num) throw la_error(LA_ERR_STATE, "preconditioner was built from a host factor");
  NumericPlan &N = *P.num;
  LA_CHECK_ARG(a_vals || N.nnzA == 0, "null values");
  if (N.n == 0) return;
  LA_HIP(hipMemsetAsync(N.err.p, 0, sizeof(int) * 4, s));
  double *fav = N.fav.as<double>();
  hipLaunchKernelGGL(k_fill_zero, dim3(2048), dim3(256), 0, s, N.nnzF, fav);
  if (N.nnzA)
    hipLaunchKernelGGL(k_scatter, dim3(static_cast<unsigned>((N.nnzA + 255) / 256)), dim3(256),
                       0, s, N.nnzA, N.a2f.as<int64_t>(), a_vals, fav);
  if (N.ic && N.additive)
    hipLaunchKernelGGL(k_shift_first, dim3((N.n + 255) / 256), dim3(256), 0, s, N.n,
                       N.fai.as<int>(), fav, N.omega);
  else if (N.ic)
    hipLaunchKernelGGL(k_scale_first, dim3((N.n + 255) / 256), dim3(256), 0, s, N.n,
                       N.fai.as<int>(), fav, 1.0 + N.omega);
  LA_HIP(hipGetLastError());
  const int L = static_cast<int>(N.level_ptr.size()) - 1;
  for (int l = 0; l < L; ++l) {
    const int k0 = N.level_ptr[l], cnt = N.level_ptr[l + 1] - k0;
    if (cnt <= 0) continue;
    if (N.ic)
      hipLaunchKernelGGL(k_ic_rows, dim3(cnt), dim3(64), 0, s, k0, N.rows.as<int>(),
                         N.fai.as<int>(), N.faj.as<int>(), N.pptr.as<int>(), N.pk.as<int>(),
                         N.pq.as<int>(), fav, N.err.as<int>(), N.additive ? 1 : 0);
    else
      hipLaunchKernelGGL(k_ilu_rows, dim3(cnt), dim3(64), 0, s, k0, N.rows.as<int>(),
                         N.fai.as<int>(), N.faj.as<int>(), N.fu.as<int>(), fav,
                         N.diag.as<double>(), N.err.as<int>());
  }
  LA_HIP(hipGetLastError());
  int bad = 0;
  LA_HIP(hipMemcpyAsync(&bad, N.err.p, sizeof(int), hipMemcpyDeviceToHost, s));
  LA_HIP(hipStreamSynchronize(s));
  if (bad)
    throw la_error(LA_ERR_NUMERIC, !N.ic        ? "ILU factorisation met a row without a diagonal"
                                   : N.additive ? "IC factorisation hit a non-positive pivot"
                                                : "IC factorisation hit a negative pivot");
  trsv_refresh(*P.lower, fav, s);
  trsv_refresh(*P.upper, fav, s);
  LA_HIP(hipStreamSynchronize(s));
}

void prec_factor(Prec &P, std::vector<int> *fai, std::vector<int> *faj, std::vector<double> *fav,
                 hipStream_t s) {
  if (!P.num) throw la_error(LA_ERR_STATE, "preconditioner was built from a host factor");
  NumericPlan &N = *P.num;
  if (fai) *fai = N.h_fai;
  if (faj) *faj = N.h_faj;
  if (fav) {
    fav->resize(static_cast<size_t>(N.nnzF));
    if (N.nnzF)
      LA_HIP(hipMemcpyAsync(fav->data(), N.fav.p, sizeof(double) * N.nnzF, hipMemcpyDeviceToHost,
                            s));
    LA_HIP(hipStreamSynchronize(s));
  }
}

}  // namespace laamd
