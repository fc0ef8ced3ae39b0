// Host-only C entry points over build_tile_layout (tiling.cpp), built by
// tests/test_tiling.py with g++ so that the tiled schedule's layout can be
// replayed on the CPU. No GPU code is involved.
#include <cstring>
#include <string>

#include "host_algos.hpp"

namespace {
laamd::TileLayout g_layout;
std::string g_why;

void copy_out(int *dst, const std::vector<int> &src) {
  if (!src.empty()) std::memcpy(dst, src.data(), sizeof(int) * src.size());
}
}  // namespace

// 1 when a layout was built (then read it with tiling_sizes / tiling_get),
// 0 when the schedule is refused (tiling_why tells why)
extern "C" int tiling_build(int n, const int *ai, const int *aj, const int *level, int lanes,
                            int max_parts, int max_rows, int max_halo) {
  g_layout = laamd::TileLayout();
  g_why.clear();
  return laamd::build_tile_layout(n, ai, aj, level, lanes, max_parts, max_rows, max_halo,
                                  g_layout, g_why)
             ? 1
             : 0;
}

extern "C" const char *tiling_why() { return g_why.c_str(); }

// nparts, ngroups, rlanes, maxe, h_max, halo list length, tcode length
extern "C" void tiling_sizes(int *v) {
  v[0] = g_layout.nparts;
  v[1] = g_layout.ngroups;
  v[2] = g_layout.rlanes;
  v[3] = g_layout.maxe;
  v[4] = g_layout.h_max;
  v[5] = static_cast<int>(g_layout.rf_col.size());
  v[6] = static_cast<int>(g_layout.tcode.size());
}

extern "C" void tiling_get(int *part_gptr, int *grp_row, int *tperm, int *tcode, int *tsrc,
                           int *rf_ptr, int *rf_col, int *rf_slot) {
  copy_out(part_gptr, g_layout.part_gptr);
  copy_out(grp_row, g_layout.grp_row);
  copy_out(tperm, g_layout.tperm);
  copy_out(tcode, g_layout.tcode);
  copy_out(tsrc, g_layout.tsrc);
  copy_out(rf_ptr, g_layout.rf_ptr);
  copy_out(rf_col, g_layout.rf_col);
  copy_out(rf_slot, g_layout.rf_slot);
}
