// Reverse Cuthill-McKee on the host with a team of threads, bit-exact with
// the serial definition: reordering/Reordering.cpp:37-136 (SerialCM), the
// components of reordering/UnionFind.cpp:82-129,217-293, the pseudo-diameter
// of reordering/Reordering.h:73-136 and the BFS of reordering/BFS.cpp:59-180.
// What lets it run in parallel without changing the result:
//  * components: the lock-free union-find of ParUnionFindRem links every
//    root below a larger index, so whatever the order of the unions a
//    component's root is its largest node; components are taken in
//    ascending root order, members ascending;
//  * BFS: level-synchronous; each level's node set is unique (the nodes at
//    that distance), and the pseudo-diameter uses only the level sizes and
//    the last level as a set (its candidate is that set's minimum
//    (degree, index));
//  * ordering: in the serial sweep a node v of level l+1 is claimed by the
//    first node of level l (in the new order) whose row holds v -- with a
//    symmetric pattern, the level-l entry of v's own row with the smallest
//    position -- and each claimer's children are sorted by (degree, index);
//    so the new level l+1 is level l+1 sorted by (position of its claimer,
//    degree, index), built per level by counting placement.
// Structurally unsymmetric patterns (where a component's BFS from one node
// need not reach every member) take the serial code, quirks included.
// The degree array is the one the pseudo-diameter search mutates, as in the
// reference (its chosen nodes and their neighbours get INT_MAX).
#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "host_algos.hpp"
#include "team.hpp"

namespace laamd {
namespace {
using team::Barrier;
using team::Team;
using team::split;

// Level-synchronous BFS; q holds the visited nodes level by level (0-based),
// level k = q[lptr[k] .. lptr[k+1]).
struct Bfs {
  Team &T;
  int n, base;
  const int *ai, *aj;
  std::vector<unsigned char> seen;
  std::vector<int> q;
  std::vector<int64_t> lptr, off;
  std::vector<std::vector<int>> loc;
  std::vector<std::vector<std::vector<int>>> bkt;  // [scanner][owner]
  Bfs(Team &team, int n_, int base_, const int *ai_, const int *aj_)
      : T(team), n(n_), base(base_), ai(ai_), aj(aj_), seen(n_), q(n_), off(team.size()),
        loc(team.size()), bkt(team.size(), std::vector<std::vector<int>>(team.size())) {}

  // From `source` (0-based). `width` is max'ed with every level size,
  // including the final empty one, and not reset (the reference's BFS
  // object keeps it). Returns false when a level reaches `shortcut`.
  // `height` = the number of non-empty levels (on a true return).
  bool run(int source, int shortcut, int &height, int &width) {
    std::memset(seen.data(), 0, seen.size());
    q[0] = source;
    seen[source] = 1;
    lptr.assign({0, 1});
    int h = 0;
    bool cut = false;
    const int nt = T.size();
    Barrier bar(T);
    std::vector<int64_t> sizes(nt);
    for (auto &row : bkt)
      for (auto &bk : row) bk.clear();
    // every member walks the levels in step and derives the same bounds
    // from the per-member counts; member 0 keeps the level table. A node's
    // `seen` byte is written only by the member owning its index range:
    // the frontier's scanners hand candidates to their owners, so no two
    // members write one cache line of `seen`.
    T.run([&](int t, int) {
      int64_t b = 0, e = 1;
      int hh = 0;
      for (;;) {
        const int64_t nf = e - b;
        ++hh;
        if (t == 0) {
          width = std::max<int64_t>(width, nf);
          h = hh;
        }
        if (nf == 0) break;
        if (nf >= shortcut) {
          if (t == 0) cut = true;
          break;
        }
        int64_t pb, pe;
        split(nf, t, nt, pb, pe);
        auto &mine_b = bkt[t];
        for (int64_t k = b + pb; k < b + pe; ++k) {
          const int u = q[k];
          for (int j = ai[u] - base; j < ai[u + 1] - base; ++j) {
            const int v = aj[j] - base;
            if (!__atomic_load_n(&seen[v], __ATOMIC_RELAXED))
              mine_b[static_cast<int64_t>(v) * nt / n].push_back(v);
          }
        }
        bar.wait();
        auto &L = loc[t];
        L.clear();
        for (int s2 = 0; s2 < nt; ++s2)
          for (int v : bkt[s2][t])
            if (!__atomic_load_n(&seen[v], __ATOMIC_RELAXED)) {
              __atomic_store_n(&seen[v], 1, __ATOMIC_RELAXED);
              L.push_back(v);
            }
        sizes[t] = static_cast<int64_t>(L.size());
        bar.wait();
        for (auto &bk : mine_b) bk.clear();  // all owners have read them
        int64_t o = e, mine = e;
        for (int s2 = 0; s2 < nt; ++s2) {
          if (s2 == t) mine = o;
          o += sizes[s2];
        }
        std::copy(L.begin(), L.end(), q.begin() + mine);
        if (t == 0) lptr.push_back(o);
        bar.wait();  // q and sizes complete before the next level reads them
        b = e;
        e = o;
      }
    });
    if (cut) {
      height = h - 1;
      return false;
    }
    lptr.pop_back();  // the final empty level
    height = h - 1;
    return true;
  }
  int64_t levels() const { return static_cast<int64_t>(lptr.size()) - 1; }
};

// reordering/Reordering.h:73-136 (PseudoDiameter) with the "first five"
// candidates; `width` carries over between the BFS runs of one round, as in
// the reference (its BFS object's member, starting at 0).
int pseudo_peripheral(Bfs &bfs, const int *ai, const int *aj, int base, std::vector<int> &deg,
                      const int *view, int nview) {
  int source = -1, best = INT_MAX;
  for (int q = 0; q < nview; ++q)
    if (deg[view[q] - base] < best) {
      best = deg[view[q] - base];
      source = view[q];
    }
  int target = -1, forward_width = 0, backward_width = INT_MAX;
  std::vector<int> last;
  while (target == -1) {
    int width = 0, diameter = 0;
    bfs.run(source - base, INT_MAX, diameter, width);
    forward_width = width;
    const int64_t L = bfs.levels();
    last.assign(bfs.q.begin() + bfs.lptr[L - 1], bfs.q.begin() + bfs.lptr[L]);
    std::vector<int> chosen;
    while (chosen.size() < 5) {
      int min_deg = INT_MAX, sel = -1;
      for (int v0 : last) {
        const int v = v0 + base, d = deg[v0];
        if (d < min_deg) {
          min_deg = d;
          sel = v;
        } else if (d == min_deg) {
          sel = std::min(sel, v);
        }
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
      if (!bfs.run(c - base, backward_width, h, width)) continue;
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

// ParUnionFindRem's unite (reordering/UnionFind.cpp:82-129): parents only
// ever move to larger indices
void unite(std::atomic<int> *P, int x, int y) {
  for (;;) {
    int px = P[x].load(std::memory_order_relaxed), py = P[y].load(std::memory_order_relaxed);
    if (px == py) return;
    if (py < px) {
      std::swap(x, y);
      std::swap(px, py);
    }
    int expect = px;
    if (x == px) {
      if (P[x].compare_exchange_strong(expect, py, std::memory_order_relaxed)) return;
      continue;
    }
    if (P[x].compare_exchange_strong(expect, py, std::memory_order_relaxed)) x = px;
  }
}

// true when every (i, j) has its (j, i); rows need not be sorted
bool pattern_symmetric_impl(Team &T, int n, int base, const int *ai, const int *aj) {
  std::atomic<bool> ok{true};
  T.run([&](int t, int k) {
    int64_t b, e;
    split(n, t, k, b, e);
    for (int64_t i = b; i < e && ok.load(std::memory_order_relaxed); ++i)
      for (int j = ai[i] - base; j < ai[i + 1] - base; ++j) {
        const int c = aj[j] - base;
        if (c < 0 || c >= n) {
          ok.store(false, std::memory_order_relaxed);
          break;
        }
        bool found = false;
        for (int q = ai[c] - base; q < ai[c + 1] - base && !found; ++q)
          found = aj[q] - base == static_cast<int>(i);
        if (!found) {
          ok.store(false, std::memory_order_relaxed);
          break;
        }
      }
  });
  return ok.load();
}

}  // namespace

bool pattern_symmetric(int n, int base, const int *ai, const int *aj) {
  Team T(n < 4096 ? 1 : host_threads());
  return pattern_symmetric_impl(T, n, base, ai, aj);
}

int host_threads() {
  if (const char *e = std::getenv("LAAMD_HOST_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) return std::min(v, 256);
  }
  const unsigned hc = std::thread::hardware_concurrency();
  return static_cast<int>(std::clamp<unsigned>(hc, 1u, 16u));
}

void rcm(int n, int base, const int *ai, const int *aj, int *iperm, int *perm, int nthreads) {
  if (n <= 0) return;
  // the parallel form needs a structurally symmetric pattern (a node's BFS
  // parents are then the level-l entries of its own row's column set);
  // anything else takes the serial definition, quirks included
  if (!pattern_symmetric(n, base, ai, aj)) {
    rcm_serial(n, base, ai, aj, iperm, perm);
    return;
  }
  // below a few thousand nodes a team costs more than it saves
  const int nt = n < 4096 ? 1 : (nthreads > 0 ? nthreads : host_threads());
  Team T(nt);
  std::vector<int> deg(n);
  std::vector<std::atomic<int>> P(n);
  T.run([&](int t, int k) {
    int64_t b, e;
    split(n, t, k, b, e);
    for (int64_t i = b; i < e; ++i) {
      deg[i] = ai[i + 1] - ai[i];
      P[i].store(static_cast<int>(i), std::memory_order_relaxed);
    }
  });
  const int64_t nnz = static_cast<int64_t>(ai[n]) - base;
  T.run([&](int t, int k) {
    // rows split by entries, so each member gets about nnz/k unions
    int64_t eb, ee;
    split(nnz, t, k, eb, ee);
    const int r0 = static_cast<int>(std::upper_bound(ai, ai + n + 1, eb + base) - ai) - 1;
    for (int i = std::max(r0, 0); i < n && ai[i] - base < ee; ++i)
      for (int64_t j = std::max<int64_t>(ai[i] - base, eb); j < std::min<int64_t>(ai[i + 1] - base, ee); ++j)
        if (aj[j] - base > i) unite(P.data(), i, aj[j] - base);  // symmetric: one orientation
  });
  std::vector<int> comp_of(n);
  T.run([&](int t, int k) {
    int64_t b, e;
    split(n, t, k, b, e);
    for (int64_t i = b; i < e; ++i) {
      int x = static_cast<int>(i);
      for (;;) {
        const int p = P[x].load(std::memory_order_relaxed);
        if (p == x) break;
        const int gp = P[p].load(std::memory_order_relaxed);
        P[x].store(gp, std::memory_order_relaxed);  // halving: gp is an ancestor too
        x = p;
      }
      comp_of[i] = x;
    }
  });
  std::vector<int> cstart, members(n);
  std::atomic<bool> one{true};
  T.run([&](int t, int k) {
    int64_t b, e;
    split(n, t, k, b, e);
    for (int64_t i = b; i < e; ++i)
      if (comp_of[i] != n - 1) {
        one.store(false, std::memory_order_relaxed);
        break;
      }
  });
  if (one.load()) {  // connected: one component, members ascending
    cstart = {0, n};
    T.run([&](int t, int k) {
      int64_t b, e;
      split(n, t, k, b, e);
      for (int64_t i = b; i < e; ++i) members[i] = static_cast<int>(i) + base;
    });
  } else {
    std::vector<int> comp_index(n, -1);
    int ncomp = 0;
    for (int i = 0; i < n; ++i)
      if (comp_of[i] == i) comp_index[i] = ncomp++;
    cstart.assign(static_cast<size_t>(ncomp) + 1, 0);
    for (int i = 0; i < n; ++i) ++cstart[comp_index[comp_of[i]] + 1];
    for (int c = 0; c < ncomp; ++c) cstart[c + 1] += cstart[c];
    std::vector<int> fillp(cstart.begin(), cstart.end() - 1);
    for (int i = 0; i < n; ++i) members[fillp[comp_index[comp_of[i]]]++] = i + base;
  }
  const int ncomp = static_cast<int>(cstart.size()) - 1;
  Bfs bfs(T, n, base, ai, aj);
  std::vector<int> pos(n), lev(n), cnt(static_cast<size_t>(n) + 1), fill(n), tmp(n), owner(n);
  for (int c = 0; c < ncomp; ++c) {
    const int off = cstart[c], size = cstart[c + 1] - cstart[c];
    if (size <= 2) {
      for (int t = 0; t < size; ++t) iperm[off + t] = members[off + t];
      continue;
    }
    const int source = pseudo_peripheral(bfs, ai, aj, base, deg, members.data() + off, size);
    int height = 0, width = 0;
    bfs.run(source - base, INT_MAX, height, width);
    const int64_t L = bfs.levels();
    const std::vector<int64_t> &lp = bfs.lptr;
    const int *q = bfs.q.data();
    T.run([&](int t, int k) {
      for (int64_t l = t; l < L; l += k)
        for (int64_t i = lp[l]; i < lp[l + 1]; ++i) lev[q[i]] = static_cast<int>(l);
    });
    int *out = iperm + off;  // the component's order, 0-based until the end
    out[0] = source - base;
    pos[source - base] = 0;
    Barrier bar(T);
    T.run([&](int t, int k) {
      for (int64_t l = 0; l + 1 < L; ++l) {
        const int64_t pb = lp[l], pn = lp[l + 1] - lp[l];  // level l: out[pb .. pb+pn)
        const int64_t sb = lp[l + 1], sn = lp[l + 2] - lp[l + 1];
        int64_t b, e;
        // the claimer of each node v of level l+1: the first node of level l
        // (in the new order) whose row holds v -- for a symmetric pattern,
        // the level-l entry of v's own row with the smallest position
        split(sn, t, k, b, e);
        for (int64_t i = b; i < e; ++i) {
          const int v = q[sb + i];
          int o = INT_MAX;
          for (int j = ai[v] - base; j < ai[v + 1] - base; ++j) {
            const int u = aj[j] - base;
            if (lev[u] == l) o = std::min(o, pos[u]);
          }
          owner[i] = o - static_cast<int>(pb);
        }
        bar.wait();
        if (t == 0) {  // counting placement by claimer
          std::fill(cnt.begin(), cnt.begin() + pn + 1, 0);
          for (int64_t i = 0; i < sn; ++i) ++cnt[owner[i] + 1];
          for (int64_t r = 0; r < pn; ++r) cnt[r + 1] += cnt[r];
          std::copy(cnt.begin(), cnt.begin() + pn, fill.begin());
          for (int64_t i = 0; i < sn; ++i) tmp[fill[owner[i]]++] = q[sb + i];
        }
        bar.wait();
        // each claimer's children by (degree, index), then into the order
        split(pn, t, k, b, e);
        for (int64_t r = b; r < e; ++r) {
          int *g0 = tmp.data() + cnt[r], *g1 = tmp.data() + cnt[r + 1];
          std::sort(g0, g1, [&](int a2, int z) {
            return deg[a2] == deg[z] ? a2 < z : deg[a2] < deg[z];
          });
          for (int *g = g0; g != g1; ++g) {
            const int64_t p = sb + (g - tmp.data());
            out[p] = *g;
            pos[*g] = static_cast<int>(p);
          }
        }
        bar.wait();
      }
    });
    T.run([&](int t, int k) {
      int64_t b, e;
      split(size, t, k, b, e);
      for (int64_t i = b; i < e; ++i) out[i] += base;
    });
  }
  std::reverse(iperm, iperm + n);
  T.run([&](int t, int k) {
    int64_t b, e;
    split(n, t, k, b, e);
    for (int64_t i = b; i < e; ++i) perm[iperm[i] - base] = static_cast<int>(i) + base;
  });
}

}  // namespace laamd
