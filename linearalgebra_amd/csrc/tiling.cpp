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
//
// Rows longer than one lane takes (a 27-point stencil's ILU(1) factor: up
// to 31 entries) come from wider stencils, whose BFS metric is not the
// grid's. For those, a natural-order box grid (row = i + nx (j + ny k)) is
// recognised from the row indices instead, and a tile is a set of
// kBoxTileU x kBoxTileK lines along i: the factor's levels grow by one
// along i and by more along j and k, so a tile's rows of one level are at
// most one per line. A row (i, j, k) of the lower triangle reads lines
// (j - 1, k), (j - 2, k) and (j - 1 .. j + 2, k - 1), so in the skewed
// coordinate u = j + 2k every line it reads has u' <= u and k' <= k:
// tiles cut in (u, k) read only from tiles at lower (u, k) (higher for an
// upper triangle), and handing them out by kBoxTileU * (u block) +
// kBoxTileK * (k block) (reversed for an upper triangle) gives every tile
// a later ticket than the tiles it reads from, so the solve drains
// whatever number of tiles is resident. (Tiles cut in (j, k) would read
// their neighbour at higher j in the plane below, which reads back from
// them.)
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

// Box grids: every entry within kBoxReach steps of its row along each axis
constexpr int kBoxReach = 2;
constexpr int kBoxTileU = 8, kBoxTileK = 2;  // kBoxTileU * kBoxTileK = 64 / kTileRowLanes

bool box_ok(int n, const int *ai, const int *aj, int64_t nx, int64_t ny) {
  const int64_t P = nx * ny;
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int r = 0; r < n; ++r) {
    const int64_t ri = r % nx, rj = (r / nx) % ny, rk = r / P;
    for (int q = ai[r]; q < ai[r + 1]; ++q) {
      const int64_t c = aj[q];
      const int64_t ci = c % nx, cj = (c / nx) % ny, ck = c / P;
      bad |= std::llabs(ci - ri) > kBoxReach || std::llabs(cj - rj) > kBoxReach ||
             std::llabs(ck - rk) > kBoxReach;
    }
  }
  return !bad;
}

// smallest |row - column| above `above` (INT64_MAX if none)
int64_t min_offset_above(int n, const int *ai, const int *aj, int64_t above) {
  int64_t m = INT64_MAX;
#pragma omp parallel for schedule(static) reduction(min : m)
  for (int r = 0; r < n; ++r)
    for (int q = ai[r]; q < ai[r + 1]; ++q) {
      const int64_t d = std::llabs(static_cast<int64_t>(aj[q]) - r);
      if (d > above && d < m) m = d;
    }
  return m;
}

// nx, ny, nz of a natural-order box grid holding the pattern: the smallest
// offset past the reach along i is a line minus at most kBoxReach, the
// smallest past the reach along j a plane give or take kBoxReach lines and
// steps; each candidate is checked against every entry.
bool detect_box_grid(int n, const int *ai, const int *aj, int &nx, int &ny, int &nz) {
  constexpr int R = kBoxReach;
  const int64_t d1 = min_offset_above(n, ai, aj, 2 * R);
  if (d1 == INT64_MAX) return false;
  for (int64_t cx = d1; cx <= d1 + R; ++cx) {
    if (cx < 2 * R + 1 || n % cx) continue;
    const int64_t d2 = min_offset_above(n, ai, aj, R * cx + R);
    if (d2 == INT64_MAX) continue;
    for (int a = -R; a <= R; ++a)
      for (int b = -R; b <= R; ++b) {
        const int64_t P = d2 + a * cx + b;
        if (P <= 0 || P % cx || n % P) continue;
        const int64_t cy = P / cx, cz = n / P;
        if (cy < 2 * R + 1 || cz < 2 * R + 1) continue;
        if (box_ok(n, ai, aj, cx, cy)) {
          nx = static_cast<int>(cx);
          ny = static_cast<int>(cy);
          nz = static_cast<int>(cz);
          return true;
        }
      }
  }
  return false;
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

// Tiles from two axes recovered by BFS (nearest-neighbour stencils, any
// row order): part[i] per row, the tile edge and the tile count.
bool bfs_tiles(int n, const int *ai, const int *aj, int max_parts, int max_rows,
               std::vector<int> &part, int &tau, int &nparts, std::string &why) {
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
  part.assign(static_cast<size_t>(n), 0);
  tau = 0;
  nparts = 0;
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
  return true;
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
  const int64_t nnz = ai[n];
  std::vector<int> part, box_key;
  int tau = 0, nparts = 0, rl = 1;
  bool box = false;
  if (maxe <= kTileMaxEntries) {
    if (!bfs_tiles(n, ai, aj, max_parts, max_rows, part, tau, nparts, why)) return false;
  } else if (maxe <= kTileRowLanes * kTileMaxEntries) {
    int nx = 0, ny = 0, nz = 0;
    if (!detect_box_grid(n, ai, aj, nx, ny, nz)) {
      why = "rows longer than one lane takes, and no natural-order box grid";
      return false;
    }
    static_assert(kBoxTileU * kBoxTileK * kTileRowLanes == 64, "one group per level");
    // (u, k) blocks, u = j + 2k; ids of the blocks that hold rows
    const int nu = (ny - 1 + 2 * (nz - 1)) / kBoxTileU + 1;
    const int nkb = (nz - 1) / kBoxTileK + 1;
    const int64_t P = static_cast<int64_t>(nx) * ny;
    std::vector<int> raw(static_cast<size_t>(n)), id(static_cast<size_t>(nu) * nkb, -1);
    for (int r = 0; r < n; ++r) {
      const int j = static_cast<int>((r / nx) % ny), k = static_cast<int>(r / P);
      raw[r] = (k / kBoxTileK) * nu + (j + 2 * k) / kBoxTileU;
    }
    nparts = 0;
    for (int r = 0; r < n; ++r)
      if (id[raw[r]] < 0) id[raw[r]] = nparts++;
    part.resize(static_cast<size_t>(n));
    box_key.assign(static_cast<size_t>(nparts), 0);
    for (int r = 0; r < n; ++r) {
      part[r] = id[raw[r]];
      box_key[part[r]] = kBoxTileU * (raw[r] % nu) + kBoxTileK * (raw[r] / nu);
    }
    tau = kBoxTileU;
    rl = kTileRowLanes;
    box = true;
  } else {
    why = "rows longer than the tiled kernel takes";
    return false;
  }
  const int rows_per_group = lanes / rl;
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
  if (box) {
    // by kBoxTileU * (u block) + kBoxTileK * (k block), descending for an
    // upper triangle
    bool upper = false;
    for (int i = 0; i < n; ++i)
      if (ai[i] < ai[i + 1]) {
        upper = aj[ai[i]] > i;
        break;
      }
    std::stable_sort(porder.begin(), porder.end(), [&](int a, int b) {
      return upper ? box_key[a] > box_key[b] : box_key[a] < box_key[b];
    });
  } else {
    std::stable_sort(porder.begin(), porder.end(), [&](int a, int b) {
      return level[order[pstart[a]]] < level[order[pstart[b]]];
    });
  }
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
      // a new group at a new level or after rows_per_group rows
      const int gs = out.grp_row.back();
      if (k > prow0[p] && (level[r] != level[out.tperm[k - 1]] || k - gs >= rows_per_group))
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
  // entry codes in tiled order, rl slots of maxe entries per row: a tile's
  // own earlier rows live in a window of the last kTileWindow groups, row
  // w of the tile's group i in slot (i % kTileWindow) * lanes + rl * w
  // (the first of its lanes); a value of another tile in halo slot h is
  // -(h+1); padding reads the zero slot kTileWindow * lanes. Also the
  // source entry of each (row entries keep their stored order, the first
  // maxe on the row's first lane) and the halo fetch lists.
  out.rlanes = rl;
  out.maxe = rl == 1 && maxe <= 4 ? 4 : kTileMaxEntries;
  const int E = out.maxe;
  const int zero_slot = kTileWindow * lanes;
  std::vector<int> slot_of(static_cast<size_t>(n)), grp_of(static_cast<size_t>(n));
  for (int pi = 0; pi < nparts; ++pi)
    for (int g = out.part_gptr[pi]; g < out.part_gptr[pi + 1]; ++g)
      for (int kk = out.grp_row[g]; kk < out.grp_row[g + 1]; ++kk) {
        grp_of[out.tperm[kk]] = g - out.part_gptr[pi];
        slot_of[out.tperm[kk]] =
            ((g - out.part_gptr[pi]) % kTileWindow) * lanes + rl * (kk - out.grp_row[g]);
      }
  out.tcode.assign(static_cast<size_t>(n) * rl * E, zero_slot);
  out.tsrc.assign(static_cast<size_t>(n) * rl * E, -1);
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
          const size_t te = (static_cast<size_t>(kk) * rl + e / E) * E + e % E;
          out.tcode[te] = code;
          out.tsrc[te] = q;
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
  // (box tiles are narrow: most of their reads cross tiles, but each
  // value is fetched once and the halo is bounded above)
  if (!box && out.remote_frac > 0.35) {
    why = "too many values cross tiles";
    return false;
  }
  return true;
}

}  // namespace laamd
