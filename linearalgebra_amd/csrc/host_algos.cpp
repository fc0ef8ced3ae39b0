// Host-side symbolic work of the hot path: level sets, L/D/U splits,
// transpose, permutation, RCM, ILU(k)/IC(k) symbolic and numeric phases.
// Everything here is serial, deterministic and bit-exact with the
// reference's definitions (cited per function); none of it touches the GPU.
#include "host_algos.hpp"
#include "team.hpp"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace laamd {

// ----------------------------------------------------------- level sets
// sparse_mat_op/matrix_utils.hpp:637-675 (TopologicalSort2): depth of row i
// is 1 + the deepest row it references; rows are counting-sorted by depth,
// ascending index inside a level. The input must be strictly triangular.
int topo_sort2(int n, int base, const int *ai, const int *aj, bool upper,
               int *iperm, int *prefix) {
  std::vector<int> depth(n, 0);
  int levels = 0;
  auto visit = [&](int i) {
    int d = depth[i];
    for (int j = ai[i] - base; j < ai[i + 1] - base; ++j)
      d = std::max(d, depth[aj[j] - base] + 1);
    depth[i] = d;
    levels = std::max(levels, d + 1);
  };
  if (upper)
    for (int i = n - 1; i >= 0; --i) visit(i);
  else
    for (int i = 0; i < n; ++i) visit(i);
  std::vector<int> start(levels + 1, 0);
  for (int i = 0; i < n; ++i) ++start[depth[i] + 1];
  for (int l = 0; l < levels; ++l) start[l + 1] += start[l];
  for (int l = 0; l <= levels; ++l) prefix[l] = start[l];
  for (int i = 0; i < n; ++i) iperm[start[depth[i]]++] = i + base;
  return levels;
}

// ----------------------------------------------------------- splits
// sparse_mat_op/matrix_utils.hpp:720-803 (SplitLDU). A missing diagonal
// gives D[i] = 0 and leaves every entry of the row in L or U.
void split_ldu(int n, int base, const int *ai, const int *aj, const double *av,
               int *Lai, int *Laj, double *Lav, double *D, int *Uai, int *Uaj,
               double *Uav, int64_t *lnnz, int64_t *unnz) {
  int64_t lp = 0, up = 0;
  Lai[0] = base;
  Uai[0] = base;
  for (int i = 0; i < n; ++i) {
    const int *rb = aj + (ai[i] - base), *re = aj + (ai[i + 1] - base);
    const int *mid = std::lower_bound(rb, re, i + base);
    const bool zero_diag = (mid == re || *mid != i + base);
    if (D) D[i] = zero_diag ? 0.0 : av[mid - aj];
    for (const int *p = rb; p < mid; ++p, ++lp)
      if (Laj) { Laj[lp] = *p; Lav[lp] = av[p - aj]; }
    for (const int *p = mid + (zero_diag ? 0 : 1); p < re; ++p, ++up)
      if (Uaj) { Uaj[up] = *p; Uav[up] = av[p - aj]; }
    Lai[i + 1] = static_cast<int>(lp + base);
    Uai[i + 1] = static_cast<int>(up + base);
  }
  if (lnnz) *lnnz = lp;
  if (unnz) *unnz = up;
}

// sparse_mat_op/matrix_utils.hpp:805-879 (SplitTriangle, diagonal kept).
void split_triangle(int n, int base, const int *ai, const int *aj,
                    const double *av, bool upper, int *Tai, int *Taj,
                    double *Tav, int64_t *tnnz) {
  int64_t p = 0;
  Tai[0] = base;
  for (int i = 0; i < n; ++i) {
    const int *rb = aj + (ai[i] - base), *re = aj + (ai[i + 1] - base);
    const int *s = upper ? std::lower_bound(rb, re, i + base) : rb;
    const int *e = upper ? re : std::upper_bound(rb, re, i + base);
    for (const int *q = s; q < e; ++q, ++p)
      if (Taj) { Taj[p] = *q; if (Tav) Tav[p] = av[q - aj]; }
    Tai[i + 1] = static_cast<int>(p + base);
  }
  if (tnnz) *tnnz = p;
}

// mkl_wrapper/mkl_sparse_mat.cpp:517-533 (rowwiseSqrtNorm): d_i =
// 1 / sqrt(sqrt(sum_j a_ij^2)), or 1 when that sum is not above DBL_MIN;
// the sum runs left to right over the row.
void rowwise_sqrt_norm(int n, int base, const int *ai, const double *av, double *d) {
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t j = ai[i] - base; j < ai[i + 1] - base; ++j) s += av[j] * av[j];
    d[i] = s > std::numeric_limits<double>::min() ? 1.0 / std::sqrt(std::sqrt(s)) : 1.0;
  }
}

// mkl_wrapper/mkl_sparse_mat.cpp:535-544 (DtAD): a_ij *= d_i * d_j.
void dtad(int n, int base, const int *ai, const int *aj, double *av, const double *d) {
  for (int i = 0; i < n; ++i)
    for (int64_t j = ai[i] - base; j < ai[i + 1] - base; ++j) av[j] *= d[i] * d[aj[j] - base];
}

// sparse_mat_op/matrix_utils.hpp:677-704 (Diagonal).
bool diagonal(int n, int base, const int *ai, const int *aj, const double *av,
              int *diagpos, double *diag) {
  bool ok = true;
  for (int i = 0; i < n; ++i) {
    const int *rb = aj + (ai[i] - base), *re = aj + (ai[i + 1] - base);
    const int *mid = std::lower_bound(rb, re, i + base);
    const bool found = mid != re && *mid == i + base;
    ok = ok && found;
    if (diagpos) diagpos[i] = static_cast<int>(mid - aj) + base;
    if (diag) diag[i] = found ? av[mid - aj] : 0.0;
  }
  return ok;
}

// sparse_mat_op/matrix_utils.hpp:311-390 (ParallelTranspose2): stable, so
// each transposed row lists its columns in ascending source-row order.
void transpose(int rows, int cols, int base, const int *ai, const int *aj,
               const double *av, int *ti, int *tj, double *tv) {
  std::vector<int64_t> cnt(static_cast<size_t>(cols) + 1, 0);
  const int64_t nnz = static_cast<int64_t>(ai[rows]) - base;
  for (int64_t k = 0; k < nnz; ++k) ++cnt[aj[k] - base + 1];
  for (int c = 0; c < cols; ++c) cnt[c + 1] += cnt[c];
  for (int c = 0; c <= cols; ++c) ti[c] = static_cast<int>(cnt[c] + base);
  for (int r = 0; r < rows; ++r)
    for (int64_t k = ai[r] - base; k < ai[r + 1] - base; ++k) {
      const int64_t pos = cnt[aj[k] - base]++;
      tj[pos] = r + base;
      if (tv && av) tv[pos] = av[k];
    }
}

// sparse_mat_op/matrix_utils.hpp:398-472 (permutedAI + permute): row i of
// the result is row iperm[i] of A, column c becomes perm[c]; each row is
// re-sorted by column (columns are distinct, so any stable sort agrees with
// the reference's insertion sort).
void permute(int n, int base, const int *ai, const int *aj, const double *av,
             const int *iperm, const int *perm, int *pai, int *paj,
             double *pav) {
  pai[0] = base;
  for (int i = 0; i < n; ++i) {
    const int k = iperm ? iperm[i] - base : i;
    pai[i + 1] = pai[i] + (ai[k + 1] - ai[k]);
  }
  std::vector<std::pair<int, double>> row;
  for (int i = 0; i < n; ++i) {
    const int k = iperm ? iperm[i] - base : i;
    const int64_t src = ai[k] - base, len = ai[k + 1] - ai[k];
    const int64_t dst = pai[i] - base;
    if (!perm) {
      std::copy(aj + src, aj + src + len, paj + dst);
      std::copy(av + src, av + src + len, pav + dst);
      continue;
    }
    row.resize(static_cast<size_t>(len));
    for (int64_t t = 0; t < len; ++t) row[t] = {perm[aj[src + t] - base], av[src + t]};
    std::sort(row.begin(), row.end(),
              [](const auto &a, const auto &b) { return a.first < b.first; });
    for (int64_t t = 0; t < len; ++t) {
      paj[dst + t] = row[t].first;
      pav[dst + t] = row[t].second;
    }
  }
}

// ----------------------------------------------------------- RCM
namespace {

// Frontier sizes of a BFS from `source` (reordering/BFS.cpp:59-180 as
// PBFS_Fn<true,false>): `width` is max'ed with every frontier size,
// including the final empty one, and is NOT reset (it is a member of the
// BFS object the reference reuses). Returns false when a frontier reaches
// `shortcut`. `last` receives the last non-empty level.
struct FrontierBfs {
  std::vector<unsigned char> seen;
  std::vector<int> a, b;
  FrontierBfs(int n) : seen(n), a(n), b(n) {}
  bool run(int base, const int *ai, const int *aj, int source, int shortcut,
           int &height, int &width, std::vector<int> &last) {
    std::fill(seen.begin(), seen.end(), 0);
    int *front = a.data(), *nxt = b.data(), *prev = nullptr;
    int nf = 1, nprev = 0, h = 0;
    front[0] = source - base;
    seen[source - base] = 1;
    for (;;) {
      width = std::max(width, nf);
      ++h;
      if (nf == 0) {
        last.assign(prev, prev + nprev);
        for (int &v : last) v += base;
        break;
      }
      if (nf >= shortcut) {
        height = h - 1;
        return false;
      }
      int nn = 0;
      for (int q = 0; q < nf; ++q) {
        const int u = front[q];
        for (int k = ai[u] - base; k < ai[u + 1] - base; ++k) {
          const int v = aj[k] - base;
          if (!seen[v]) { seen[v] = 1; nxt[nn++] = v; }
        }
      }
      prev = front;
      nprev = nf;
      std::swap(front, nxt);
      nf = nn;
    }
    height = h - 1;
    return true;
  }
};

// reordering/Reordering.h:73-136 (PseudoDiameter) with the "first five"
// candidates; mutates `deg` exactly as the reference does (chosen nodes and
// their neighbours become INT_MAX), which the caller's child sort then sees.
// The BFS object's width starts at 0 (the reference leaves it
// uninitialised, reordering/BFS.h:47).
int pseudo_peripheral(int n, int base, const int *ai, const int *aj,
                      std::vector<int> &deg, const int *view, int nview) {
  int source = -1, best = INT_MAX;
  for (int q = 0; q < nview; ++q)
    if (deg[view[q] - base] < best) { best = deg[view[q] - base]; source = view[q]; }
  int target = -1, forward_width = 0, backward_width = INT_MAX;
  FrontierBfs bfs(n);
  std::vector<int> last, scratch;
  while (target == -1) {
    int width = 0, diameter = 0;
    bfs.run(base, ai, aj, source, INT_MAX, diameter, width, last);
    forward_width = width;
    std::vector<int> chosen;
    while (chosen.size() < 5) {
      int min_deg = INT_MAX, sel = -1;
      for (int v : last) {
        const int d = deg[v - base];
        if (d < min_deg) { min_deg = d; sel = v; }
        else if (d == min_deg) sel = std::min(sel, v);
      }
      if (min_deg == INT_MAX) break;
      chosen.push_back(sel);
      deg[sel - base] = INT_MAX;
      for (int k = ai[sel - base] - base; k < ai[sel - base + 1] - base; ++k)
        deg[aj[k] - base] = INT_MAX;
    }
    backward_width = INT_MAX;
    for (int c : chosen) {
      int h = 0;
      if (!bfs.run(base, ai, aj, c, backward_width, h, width, scratch)) continue;
      if (diameter < h && width < backward_width) {
        source = c;
        break;
      } else if (width < backward_width) {
        backward_width = width;
        target = c;
      }
    }
    if (chosen.empty()) break;  // the reference would spin here forever
  }
  if (forward_width > backward_width) std::swap(source, target);
  return source;
}

}  // namespace

// reordering/Reordering.cpp:37-136 (SerialCM) with the components of
// reordering/UnionFind.cpp:82-129,217-293: every union-find parent points to
// a larger index, so a component's root is its largest node; components are
// taken in ascending root order with members ascending. Each component is
// ordered by BFS from a pseudo-peripheral node, children of a node sorted
// by (degree, index); the whole order is then reversed.
void rcm_serial(int n, int base, const int *ai, const int *aj, int *iperm, int *perm) {
  std::vector<int> deg(n), parent(n);
  for (int i = 0; i < n; ++i) { deg[i] = ai[i + 1] - ai[i]; parent[i] = i; }
  for (int i = 0; i < n; ++i)
    for (int k = ai[i] - base; k < ai[i + 1] - base; ++k) {
      int x = i, y = aj[k] - base;
      for (;;) {
        int px = parent[x], py = parent[y];
        if (px == py) break;
        if (py < px) { std::swap(x, y); std::swap(px, py); }
        parent[x] = py;
        if (x == px) break;
        x = px;
      }
    }
  auto find = [&](int x) {
    while (x != parent[x]) { parent[x] = parent[parent[x]]; x = parent[x]; }
    return x;
  };
  std::vector<int> comp_of(n), comp_index(n, -1);
  int ncomp = 0;
  for (int i = 0; i < n; ++i) comp_of[i] = find(i);
  for (int i = 0; i < n; ++i) if (comp_of[i] == i) comp_index[i] = ncomp++;
  std::vector<int> cstart(ncomp + 1, 0);
  for (int i = 0; i < n; ++i) ++cstart[comp_index[comp_of[i]] + 1];
  for (int c = 0; c < ncomp; ++c) cstart[c + 1] += cstart[c];
  std::vector<int> fillp(cstart.begin(), cstart.end() - 1), members(n);
  for (int i = 0; i < n; ++i) members[fillp[comp_index[comp_of[i]]]++] = i + base;

  std::vector<int> levels(n), count, children;
  for (int c = 0; c < ncomp; ++c) {
    const int off = cstart[c], size = cstart[c + 1] - cstart[c];
    if (size <= 2) {
      for (int t = 0; t < size; ++t) iperm[off + t] = members[off + t];
      continue;
    }
    const int source = pseudo_peripheral(n, base, ai, aj, deg, members.data() + off, size);
    // serial BFS levels from the source (reordering/BFS.cpp:10-57)
    std::fill(levels.begin(), levels.end(), -1);
    {
      std::vector<int> queue;
      queue.reserve(size);
      queue.push_back(source - base);
      levels[source - base] = 0;
      for (size_t h = 0; h < queue.size(); ++h) {
        const int u = queue[h];
        for (int k = ai[u] - base; k < ai[u + 1] - base; ++k) {
          const int v = aj[k] - base;
          if (levels[v] == -1) { levels[v] = levels[u] + 1; queue.push_back(v); }
        }
      }
    }
    int height = 0;
    for (int t = 0; t < size; ++t) height = std::max(height, levels[members[off + t] - base] + 1);
    count.assign(height + 1, 0);
    for (int t = 0; t < size; ++t) ++count[levels[members[off + t] - base] + 1];
    for (int l = 0; l < height; ++l) count[l + 1] += count[l];
    int e = off;
    iperm[e++] = source;
    for (int l = 0; l < height; ++l)
      for (int r = count[l]; r != count[l + 1]; ++r) {
        children.clear();
        const int u = iperm[r + off] - base;
        for (int k = ai[u] - base; k != ai[u + 1] - base; ++k) {
          const int v = aj[k] - base;
          if (levels[v] == l + 1) { children.push_back(v); levels[v] = -1; }
        }
        std::sort(children.begin(), children.end(), [&](int a, int b) {
          return deg[a] == deg[b] ? a < b : deg[a] < deg[b];
        });
        for (int v : children) iperm[e++] = v + base;
      }
  }
  std::reverse(iperm, iperm + n);
  for (int i = 0; i < n; ++i) perm[iperm[i] - base] = i + base;
}

// ----------------------------------------------------------- ILU(k)
// mkl_wrapper/incomplete_lu.cpp:153-275. Row i starts as A's row (level 0);
// for each k < i in the growing row, row k's upper part U(k, j>k) is merged
// with level lev_ik + lev_kj + 1, kept if <= level, min on collision. The
// row lives in a sorted linked list over a dense column index (next[], HEAD
// = n), which is the reference's forward_list laid out in arrays.
// first_upper is produced for level 0 as well (see the numeric phase).
namespace {
// ILU(0) / ILU(1) symbolic with a team of threads, equal to the serial
// rule below. A level-1 entry of lev_ij = min(lev_ij, lev_ik + lev_kj + 1)
// can only come from lev_ik = lev_kj = 0, i.e. from entries of A: (i, j) is
// in the ILU(1) pattern iff it is in A or some k < i has (i, k) in A and
// (k, j) in A with j > k. Each row is then independent. Needs rows with
// strictly increasing columns (the serial code's merge assumes it too).
bool rows_strictly_sorted(team::Team &T, int n, int base, const int *ai, const int *aj) {
  std::atomic<bool> ok{true};
  T.run([&](int t, int k) {
    int64_t b, e;
    team::split(n, t, k, b, e);
    for (int64_t i = b; i < e && ok.load(std::memory_order_relaxed); ++i)
      for (int q = ai[i] - base + 1; q < ai[i + 1] - base; ++q)
        if (aj[q] <= aj[q - 1]) {
          ok.store(false, std::memory_order_relaxed);
          break;
        }
  });
  return ok.load();
}

void iluk_symbolic_le1(team::Team &T, int n, int base, const int *ai, const int *aj, int level,
                       std::vector<int> &fai, std::vector<int> &faj,
                       std::vector<int> &first_upper) {
  const int nt = T.size();
  std::vector<std::vector<int>> out(nt), len(nt);
  T.run([&](int t, int k) {
    int64_t rb, re;
    team::split(n, t, k, rb, re);
    auto &o = out[t];
    auto &l = len[t];
    l.assign(static_cast<size_t>(re - rb), 0);
    std::vector<int> mark(level > 0 ? n : 0, -1), cand;
    for (int64_t i = rb; i < re; ++i) {
      const size_t start = o.size();
      if (level == 0) {
        for (int q = ai[i] - base; q < ai[i + 1] - base; ++q) o.push_back(aj[q]);
      } else {
        cand.clear();
        for (int q = ai[i] - base; q < ai[i + 1] - base; ++q) {
          const int c = aj[q] - base;
          mark[c] = static_cast<int>(i);
          cand.push_back(c);
        }
        for (int q = ai[i] - base; q < ai[i + 1] - base; ++q) {
          const int m = aj[q] - base;
          if (m >= i) break;  // sorted: the rest are not below i
          for (int r = ai[m] - base; r < ai[m + 1] - base; ++r) {
            const int c = aj[r] - base;
            if (c <= m || mark[c] == i) continue;
            mark[c] = static_cast<int>(i);
            cand.push_back(c);
          }
        }
        std::sort(cand.begin(), cand.end());
        for (int c : cand) o.push_back(c + base);
      }
      l[i - rb] = static_cast<int>(o.size() - start);
    }
  });
  fai.assign(static_cast<size_t>(n) + 1, base);
  std::vector<int64_t> off(static_cast<size_t>(nt) + 1, 0);
  for (int t = 0; t < nt; ++t) off[t + 1] = off[t] + static_cast<int64_t>(out[t].size());
  if (off[nt] + base > INT_MAX) throw std::length_error("ILU pattern exceeds int32 indices");
  faj.resize(static_cast<size_t>(off[nt]));
  first_upper.assign(n, 0);
  // row pointers: per member, then shifted by the members before it
  T.run([&](int t, int k) {
    int64_t rb, re;
    team::split(n, t, k, rb, re);
    int64_t p = off[t];
    for (int64_t i = rb; i < re; ++i) {
      const int64_t rs = p;
      p += len[t][i - rb];
      fai[i + 1] = static_cast<int>(p) + base;
      int64_t fu = p;  // first column above i, else the row's end
      for (int64_t q = rs; q < p; ++q)
        if (out[t][q - off[t]] - base > i) {
          fu = q;
          break;
        }
      first_upper[i] = static_cast<int>(fu);
    }
    std::copy(out[t].begin(), out[t].end(), faj.begin() + off[t]);
  });
}
}  // namespace

void iluk_symbolic(int n, int base, const int *ai, const int *aj, int level,
                   std::vector<int> &fai, std::vector<int> &faj,
                   std::vector<int> &first_upper) {
  if (level <= 1 && n >= 4096) {
    team::Team T(host_threads());
    if (rows_strictly_sorted(T, n, base, ai, aj)) {
      iluk_symbolic_le1(T, n, base, ai, aj, level, fai, faj, first_upper);
      return;
    }
  }
  fai.assign(static_cast<size_t>(n) + 1, base);
  first_upper.assign(n, 0);
  faj.clear();
  faj.reserve(static_cast<size_t>(ai[n] - base) * (level > 0 ? 2 : 1));
  std::vector<int> flev;  // level of each stored pattern entry
  flev.reserve(faj.capacity());
  std::vector<int> next(static_cast<size_t>(n) + 1), lev(n);
  const int HEAD = n, END = -1;
  for (int i = 0; i < n; ++i) {
    int prev = HEAD;
    for (int k = ai[i] - base; k < ai[i + 1] - base; ++k) {
      const int c = aj[k] - base;
      next[prev] = c;
      lev[c] = 0;
      prev = c;
    }
    next[prev] = END;
    if (level > 0) {
      for (int k = next[HEAD]; k != END && k < i; k = next[k]) {
        const int lik = lev[k];
        if (lik >= level) continue;  // lik + lev_kj + 1 > level for every j
        int e = k;
        for (int64_t j = first_upper[k]; j < fai[k + 1] - base; ++j) {
          const int nl = lik + flev[j] + 1;
          if (nl > level) continue;
          const int c = faj[j] - base;
          while (next[e] != END && next[e] <= c) e = next[e];
          if (e == c) {
            if (lev[e] > nl) lev[e] = nl;
          } else {
            next[c] = next[e];
            next[e] = c;
            lev[c] = nl;
            e = c;
          }
        }
      }
    }
    const int64_t row_start = static_cast<int64_t>(faj.size());
    int64_t fu = -1;
    for (int c = next[HEAD]; c != END; c = next[c]) {
      if (fu < 0 && c > i) fu = static_cast<int64_t>(faj.size());
      faj.push_back(c + base);
      flev.push_back(lev[c]);
    }
    (void)row_start;
    fai[i + 1] = static_cast<int>(faj.size()) + base;
    first_upper[i] = static_cast<int>(fu < 0 ? static_cast<int64_t>(faj.size()) : fu);
  }
}

// mkl_wrapper/incomplete_lu.cpp:277-325: scatter A into the pattern, then
// row-IKJ elimination: l_ik = a_ik / u_kk, a_ij -= l_ik * u_kj over the
// merged columns j > k. No pivoting. Returns false when a row has no
// diagonal entry (the reference prints "no element on diagonal!").
bool iluk_numeric(int n, int base, const int *ai, const int *aj,
                  const double *av, const int *fai, const int *faj,
                  const int *first_upper, double *fav) {
  std::vector<double> diag(n, 0.0);
  bool ok = true;
  for (int i = 0; i < n; ++i) {
    const int64_t fb = fai[i] - base, fe = fai[i + 1] - base;
    int64_t a = ai[i] - base;
    const int64_t ae = ai[i + 1] - base;
    for (int64_t k = fb; k < fe; ++k) {
      if (a == ae || faj[k] != aj[a]) fav[k] = 0.0;
      else fav[k] = av[a++];
    }
    int64_t k_idx = fb;
    for (; k_idx < fe && faj[k_idx] - base != i; ++k_idx) {
      const int k = faj[k_idx] - base;
      fav[k_idx] /= diag[k];
      const double lik = fav[k_idx];
      int64_t p = k_idx, q = first_upper[k];
      const int64_t qe = fai[k + 1] - base;
      while (q != qe && p != fe) {
        if (faj[p] == faj[q]) fav[p++] -= lik * fav[q++];
        else if (faj[p] < faj[q]) ++p;
        else ++q;
      }
    }
    if (k_idx == fe) ok = false;
    else diag[i] = fav[k_idx];
  }
  return ok;
}

// ----------------------------------------------------------- IC(k)
// sparse_mat_op/precond.hpp:39-143 (ICCLevelSymbolic0; the four reference
// variants produce the same pattern, tests/precond_test.cpp:63-127). Rows of
// the upper factor are built one at a time; llist/jk thread the rows k < j
// whose next upper entry is column j (LIFO lists ending at INT_MAX).
void icc_symbolic(int n, int base, const int *ai, const int *aj,
                  const int *diag_pos, int lvl, std::vector<int> &icc_ai,
                  std::vector<int> &icc_aj) {
  const int NONE = INT_MAX, HEAD = n, END = -1;
  std::vector<int> llist(n, NONE), next(static_cast<size_t>(n) + 1), lev(n), flev;
  std::vector<int64_t> jk(n, 0);
  icc_ai.assign(static_cast<size_t>(n) + 1, base);
  icc_aj.clear();
  icc_aj.reserve(static_cast<size_t>(ai[n] - base));
  flev.reserve(icc_aj.capacity());
  for (int j = 0; j < n; ++j) {
    int prev = HEAD;
    for (int64_t p = diag_pos[j] - base; p < ai[j + 1] - base; ++p) {
      const int c = aj[p] - base;
      next[prev] = c;
      lev[c] = 0;
      prev = c;
    }
    next[prev] = END;
    for (int k = llist[j]; k < j;) {
      int64_t q = jk[k]++;
      const int64_t qe = icc_ai[k + 1] - base;
      const int knext = llist[k];
      if (q + 1 < qe) {
        const int c = icc_aj[q + 1] - base;
        llist[k] = llist[c];
        llist[c] = k;
      }
      const int lik = flev[q];
      int e = next[HEAD];
      for (; q < qe; ++q) {
        const int nl = lik + flev[q] + 1;
        if (nl > lvl) continue;
        const int c = icc_aj[q] - base;
        while (next[e] != END && next[e] <= c) e = next[e];
        if (e == c) {
          if (nl < lev[e]) lev[e] = nl;
        } else {
          next[c] = next[e];
          next[e] = c;
          lev[c] = nl;
          e = c;
        }
      }
      k = knext;
    }
    const int64_t rs = static_cast<int64_t>(icc_aj.size());
    for (int c = next[HEAD]; c != END; c = next[c]) {
      icc_aj.push_back(c + base);
      flev.push_back(lev[c]);
    }
    icc_ai[j + 1] = static_cast<int>(icc_aj.size()) + base;
    if (static_cast<int64_t>(icc_aj.size()) - rs > 1) {
      const int c = icc_aj[rs + 1] - base;
      llist[j] = llist[c];
      llist[c] = j;
      jk[j] = rs + 1;
    }
  }
}

// sparse_mat_op/precond.hpp:615-693 (ICCLevelNumeric): up-looking IC on the
// upper pattern with the diagonal shifted by (1 + omega); rows k are visited
// in the same LIFO list order as the reference, so the result is bitwise the
// same. Returns false on a negative pivot.
bool icc_numeric(int n, int base, const int *ai, const int *aj, const double *av,
                 const int *diag_pos, double omega, const int *icc_ai,
                 const int *icc_aj, double *icc_av) {
  const int NONE = INT_MAX;
  std::vector<int> llist(n, NONE);
  std::vector<int64_t> jk(n, 0);
  for (int j = 0; j < n; ++j) {
    int64_t a = diag_pos[j] - base;
    const int64_t ae = ai[j + 1] - base;
    const int64_t ps = icc_ai[j] - base, pe = icc_ai[j + 1] - base;
    int64_t p = ps;
    icc_av[p++] = av[a++] * (1.0 + omega);
    for (; p < pe; ++p) {
      if (a == ae || icc_aj[p] != aj[a]) icc_av[p] = 0.0;
      else icc_av[p] = av[a++];
    }
    for (int k = llist[j]; k < j;) {
      int64_t q = jk[k]++;
      const int64_t qe = icc_ai[k + 1] - base;
      const int knext = llist[k];
      if (q + 1 < qe) {
        const int c = icc_aj[q + 1] - base;
        llist[k] = llist[c];
        llist[c] = k;
      }
      k = knext;
      const double ukj = icc_av[q];
      int64_t t = ps;
      while (q < qe && t < pe) {
        if (icc_aj[q] == icc_aj[t]) icc_av[t++] -= ukj * icc_av[q++];
        else if (icc_aj[q] < icc_aj[t]) ++q;
        else ++t;
      }
    }
    if (pe - ps > 1) {
      const int c = icc_aj[ps + 1] - base;
      llist[j] = llist[c];
      llist[c] = j;
      jk[j] = ps + 1;
    }
    if (icc_av[ps] < 0) return false;
    const double d = std::sqrt(icc_av[ps]);
    icc_av[ps] = d;
    for (int64_t t = ps + 1; t < pe; ++t) icc_av[t] /= d;
  }
  return true;
}

// mkl_wrapper/incomplete_cholesky.cpp:195-274 (incomplete_cholesky_k::
// numeric_factorize, one pass): the same up-looking IC as icc_numeric but the
// diagonal gets an ADDITIVE shift, the rows k with an entry in column j are
// visited first-in first-out (the reference keeps one append-only vector per
// column; here a per-column queue threaded through next[]), and a pivot <= 0
// fails. The two visit orders round differently, so the factors agree only
// to a few ulps. diag_pos is base-adjusted like icc_numeric's.
bool ick_numeric(int n, int base, const int *ai, const int *aj, const double *av,
                 const int *diag_pos, double shift, const int *icc_ai,
                 const int *icc_aj, double *icc_av) {
  const int NONE = -1;
  std::vector<int> head(n, NONE), tail(n, NONE), next(n, NONE);
  std::vector<int64_t> pos(n, 0);
  auto enqueue = [&](int k, int c) {
    next[k] = NONE;
    if (head[c] == NONE) head[c] = k;
    else next[tail[c]] = k;
    tail[c] = k;
  };
  for (int j = 0; j < n; ++j) {
    int64_t a = diag_pos[j] - base;
    const int64_t ae = ai[j + 1] - base;
    const int64_t ps = icc_ai[j] - base, pe = icc_ai[j + 1] - base;
    for (int64_t p = ps; p < pe; ++p) {
      if (a == ae || icc_aj[p] != aj[a]) icc_av[p] = 0.0;
      else icc_av[p] = av[a++];
    }
    icc_av[ps] += shift;
    for (int k = head[j]; k != NONE;) {
      const int knext = next[k];
      int64_t q = pos[k];
      const int64_t qe = icc_ai[k + 1] - base;
      if (q + 1 < qe) {
        pos[k] = q + 1;
        enqueue(k, icc_aj[q + 1] - base);
      }
      k = knext;
      const double ukj = icc_av[q];
      int64_t t = ps;
      while (q < qe && t < pe) {
        if (icc_aj[q] == icc_aj[t]) icc_av[t++] -= ukj * icc_av[q++];
        else if (icc_aj[q] < icc_aj[t]) ++q;
        else ++t;
      }
    }
    head[j] = NONE;
    if (pe - ps > 1) {
      pos[j] = ps + 1;
      enqueue(j, icc_aj[ps + 1] - base);
    }
    if (icc_av[ps] <= 0) return false;
    const double d = std::sqrt(icc_av[ps]);
    icc_av[ps] = d;
    for (int64_t t = ps + 1; t < pe; ++t) icc_av[t] /= d;
  }
  return true;
}

}  // namespace laamd
