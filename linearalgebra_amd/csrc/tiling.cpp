// Tiled schedule of the sync-free triangular solve: which workgroup solves
// which rows.
//
// The level sets of a triangle from a structured grid are diagonal planes,
// and a solve that hands every level to whichever workgroup is free pays one
// cross-workgroup hop per level. Here the rows are cut into tiles that run
// along the levels instead: a tile is a column of the grid, every level of
// it is solved by the same wavefront, and a dependency inside the tile is a
// value the wavefront itself wrote into LDS one level earlier. Only the
// tile's faces read values from other tiles.
//
// The grid's coordinates are recovered from the pattern alone, whatever the
// row order (natural, RCM, any other): BFS distances from a corner s and
// from the nearest other corners t (corners are the nodes of least degree)
// give f = d(s) - d(t), which on a box grid with a nearest-neighbour stencil
// is twice the coordinate along the s-t edge, up to a shift. Two such axes
// give 2-D tiles; when the pattern does not look like a grid (no two
// independent axes, rows longer than the kernel takes, too many values
// crossing tiles, tiles too large for LDS) the layout is refused and the
// solve keeps the level-chunk schedule.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "host_algos.hpp"

namespace laamd {

namespace {

// BFS distances on the symmetric adjacency (ptr/adj); -1 where unreached
void bfs(int n, const std::vector<int> &ptr, const std::vector<int> &adj, int s,
         std::vector<int> &dist, std::vector<int> &queue) {
  dist.assign(static_cast<size_t>(n), -1);
  queue.resize(static_cast<size_t>(n));
  int head = 0, tail = 0;
  dist[s] = 0;
  queue[tail++] = s;
  while (head < tail) {
    const int u = queue[head++];
    for (int q = ptr[u]; q < ptr[u + 1]; ++q) {
      const int v = adj[q];
      if (dist[v] < 0) {
        dist[v] = dist[u] + 1;
        queue[tail++] = v;
      }
    }
  }
}

double correlation(const std::vector<int> &a, const std::vector<int> &b) {
  const size_t n = a.size();
  double ma = 0, mb = 0;
  for (size_t i = 0; i < n; ++i) {
    ma += a[i];
    mb += b[i];
  }
  ma /= n;
  mb /= n;
  double sab = 0, saa = 0, sbb = 0;
  for (size_t i = 0; i < n; ++i) {
    const double x = a[i] - ma, y = b[i] - mb;
    sab += x * y;
    saa += x * x;
    sbb += y * y;
  }
  if (saa <= 0 || sbb <= 0) return 1.0;
  return sab / std::sqrt(saa * sbb);
}

}  // namespace

bool build_tile_layout(int n, const int *ai, const int *aj, const int *level, int lanes,
                       int max_parts, int max_rows, int max_halo, TileLayout &out,
                       std::string &why) {
  out = TileLayout();
  if (n < 4 * lanes) {
    why = "too few rows";
    return false;
  }
  int maxe = 0;
  for (int i = 0; i < n; ++i) maxe = std::max(maxe, ai[i + 1] - ai[i]);
  if (maxe > kTileMaxEntries) {
    why = "rows longer than the tiled kernel takes";
    return false;
  }
  const int64_t nnz = ai[n];
  // symmetric adjacency of the triangle (T + T^T)
  std::vector<int> ptr(static_cast<size_t>(n) + 1, 0), adj(static_cast<size_t>(2 * nnz));
  for (int i = 0; i < n; ++i) {
    ptr[i + 1] += ai[i + 1] - ai[i];
    for (int q = ai[i]; q < ai[i + 1]; ++q) ptr[aj[q] + 1]++;
  }
  for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  {
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int q = ai[i]; q < ai[i + 1]; ++q) {
        adj[fill[i]++] = aj[q];
        adj[fill[aj[q]]++] = i;
      }
  }
  // corners: the nodes of least degree (at most 64 of them looked at)
  int mindeg = INT32_MAX;
  for (int i = 0; i < n; ++i) mindeg = std::min(mindeg, ptr[i + 1] - ptr[i]);
  std::vector<int> corners;
  for (int i = 0; i < n && corners.size() < 64; ++i)
    if (ptr[i + 1] - ptr[i] == mindeg) corners.push_back(i);
  if (corners.size() < 3) {
    why = "no grid corners";
    return false;
  }
  std::vector<int> ds, dt, queue;
  const int s = corners[0];
  bfs(n, ptr, adj, s, ds, queue);
  for (int i = 0; i < n; ++i)
    if (ds[i] < 0) {
      why = "pattern not connected";
      return false;
    }
  std::vector<int> cand(corners.begin() + 1, corners.end());
  std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return ds[a] < ds[b]; });
  std::vector<std::vector<int>> axes;
  for (size_t c = 0; c < cand.size() && c < 6 && axes.size() < 2; ++c) {
    bfs(n, ptr, adj, cand[c], dt, queue);
    std::vector<int> f(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) f[i] = ds[i] - dt[i];
    const auto mm = std::minmax_element(f.begin(), f.end());
    bool indep = *mm.second - *mm.first >= 4;
    for (auto &a : axes)
      if (std::fabs(correlation(a, f)) > 0.5) indep = false;
    if (indep) axes.push_back(std::move(f));
  }
  if (axes.size() < 2) {
    why = "no two independent grid axes";
    return false;
  }
  std::vector<int> u(static_cast<size_t>(n)), v(static_cast<size_t>(n));
  {
    const int m0 = *std::min_element(axes[0].begin(), axes[0].end());
    const int m1 = *std::min_element(axes[1].begin(), axes[1].end());
    for (int i = 0; i < n; ++i) {
      u[i] = (axes[0][i] - m0) / 2;
      v[i] = (axes[1][i] - m1) / 2;
    }
  }
  const int umax = *std::max_element(u.begin(), u.end());
  const int vmax = *std::max_element(v.begin(), v.end());
  // the smallest square tile whose count fits the resident workgroups and
  // whose rows fit LDS
  std::vector<int> part(static_cast<size_t>(n));
  int tau = 0, nparts = 0;
  for (int t = 2; t <= 4096; t *= 2) {
    const int nu = umax / t + 1, nv = vmax / t + 1;
    if (static_cast<int64_t>(nu) * nv > 4 * static_cast<int64_t>(max_parts)) continue;
    std::vector<int> cnt(static_cast<size_t>(nu) * nv, 0);
    for (int i = 0; i < n; ++i) cnt[static_cast<size_t>(u[i] / t) * nv + v[i] / t]++;
    int used = 0, big = 0;
    for (int c : cnt) {
      used += c > 0;
      big = std::max(big, c);
    }
    if (used <= max_parts && big <= max_rows) {
      std::vector<int> id(cnt.size(), -1);
      nparts = 0;
      for (size_t c = 0; c < cnt.size(); ++c)
        if (cnt[c]) id[c] = nparts++;
      for (int i = 0; i < n; ++i) part[i] = id[static_cast<size_t>(u[i] / t) * nv + v[i] / t];
      tau = t;
      break;
    }
  }
  if (!tau) {
    why = "no tile size fits";
    return false;
  }
  // rows of each tile by (level, index); tiles by their first level
  std::vector<int> order(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    if (part[a] != part[b]) return part[a] < part[b];
    if (level[a] != level[b]) return level[a] < level[b];
    return a < b;
  });
  std::vector<int> pstart(static_cast<size_t>(nparts) + 1, 0);
  for (int i = 0; i < n; ++i) pstart[part[i] + 1]++;
  for (int p = 0; p < nparts; ++p) pstart[p + 1] += pstart[p];
  std::vector<int> porder(static_cast<size_t>(nparts));
  for (int p = 0; p < nparts; ++p) porder[p] = p;
  std::stable_sort(porder.begin(), porder.end(), [&](int a, int b) {
    return level[order[pstart[a]]] < level[order[pstart[b]]];
  });
  // row positions: tiles in ticket order, rows by (level, index) inside
  std::vector<int> prow0(static_cast<size_t>(nparts));
  out.tperm.resize(static_cast<size_t>(n));
  out.part_gptr.assign(1, 0);
  out.grp_row.assign(1, 0);
  int k = 0;
  for (int pi = 0; pi < nparts; ++pi) {
    const int p = porder[pi];
    prow0[p] = k;
    for (int t = pstart[p]; t < pstart[p + 1]; ++t) {
      const int r = order[t];
      // a new group at a new level or after `lanes` rows
      const int gs = out.grp_row.back();
      if (k > prow0[p] && (level[r] != level[out.tperm[k - 1]] || k - gs >= lanes))
        out.grp_row.push_back(k);
      else if (k == prow0[p] && k > 0)
        out.grp_row.push_back(k);
      out.tperm[k++] = r;
    }
    out.m_max = std::max(out.m_max, pstart[p + 1] - pstart[p]);
    out.part_gptr.push_back(static_cast<int>(out.grp_row.size()));
  }
  out.grp_row.push_back(n);
  // part_gptr[pi + 1] counts the groups opened through tile pi; a group
  // ends where the next one starts
  out.nparts = nparts;
  out.ngroups = static_cast<int>(out.grp_row.size()) - 1;
  // entry codes in tiled order, maxe per row: a tile's own earlier rows
  // live in a window of the last kTileWindow groups, row `lane` of the
  // tile's group i in slot (i % kTileWindow) * lanes + lane; a value of
  // another tile in halo slot h is -(h+1); padding reads the zero slot
  // kTileWindow * lanes. Also the source entry of each (row entries keep
  // their stored order) and the halo fetch lists.
  out.maxe = maxe <= 4 ? 4 : kTileMaxEntries;
  const int E = out.maxe;
  const int zero_slot = kTileWindow * lanes;
  std::vector<int> slot_of(static_cast<size_t>(n)), grp_of(static_cast<size_t>(n));
  for (int pi = 0; pi < nparts; ++pi)
    for (int g = out.part_gptr[pi]; g < out.part_gptr[pi + 1]; ++g)
      for (int kk = out.grp_row[g]; kk < out.grp_row[g + 1]; ++kk) {
        grp_of[out.tperm[kk]] = g - out.part_gptr[pi];
        slot_of[out.tperm[kk]] =
            ((g - out.part_gptr[pi]) % kTileWindow) * lanes + (kk - out.grp_row[g]);
      }
  out.tcode.assign(static_cast<size_t>(n) * E, zero_slot);
  out.tsrc.assign(static_cast<size_t>(n) * E, -1);
  out.rf_ptr.assign(1, 0);
  std::vector<int> halo_of(static_cast<size_t>(n), -1), halo_used;
  int64_t remote = 0;
  for (int pi = 0; pi < nparts; ++pi) {
    const int p = porder[pi];
    for (int h : halo_used) halo_of[h] = -1;
    halo_used.clear();
    for (int g = out.part_gptr[pi]; g < out.part_gptr[pi + 1]; ++g) {
      const int gi = g - out.part_gptr[pi];
      for (int kk = out.grp_row[g]; kk < out.grp_row[g + 1]; ++kk) {
        const int r = out.tperm[kk];
        for (int q = ai[r], e = 0; q < ai[r + 1]; ++q, ++e) {
          const int c = aj[q];
          int code;
          if (part[c] == p) {
            if (gi - grp_of[c] >= kTileWindow || grp_of[c] >= gi) {
              why = "dependencies reach beyond the LDS window";
              return false;
            }
            code = slot_of[c];
          } else {
            ++remote;
            if (halo_of[c] < 0) {
              halo_of[c] = static_cast<int>(halo_used.size());
              halo_used.push_back(c);
              out.rf_col.push_back(c);
              out.rf_slot.push_back(halo_of[c]);
            }
            code = -(halo_of[c] + 1);
          }
          out.tcode[static_cast<size_t>(kk) * E + e] = code;
          out.tsrc[static_cast<size_t>(kk) * E + e] = q;
        }
      }
      out.rf_ptr.push_back(static_cast<int>(out.rf_col.size()));
    }
    out.h_max = std::max(out.h_max, static_cast<int>(halo_used.size()));
  }
  out.tau = tau;
  out.remote_frac = nnz ? static_cast<double>(remote) / nnz : 0.0;
  if (out.h_max > max_halo) {
    why = "tile faces too large for LDS";
    return false;
  }
  if (out.remote_frac > 0.35) {
    why = "too many values cross tiles";
    return false;
  }
  return true;
}

}  // namespace laamd
