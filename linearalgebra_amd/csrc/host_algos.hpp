// Host-side (CPU) algorithms of the hot path; see host_algos.cpp.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace laamd {

int topo_sort2(int n, int base, const int *ai, const int *aj, bool upper,
               int *iperm, int *prefix);
void split_ldu(int n, int base, const int *ai, const int *aj, const double *av,
               int *Lai, int *Laj, double *Lav, double *D, int *Uai, int *Uaj,
               double *Uav, int64_t *lnnz, int64_t *unnz);
void split_triangle(int n, int base, const int *ai, const int *aj,
                    const double *av, bool upper, int *Tai, int *Taj,
                    double *Tav, int64_t *tnnz);
void rowwise_sqrt_norm(int n, int base, const int *ai, const double *av, double *d);
void dtad(int n, int base, const int *ai, const int *aj, double *av, const double *d);
bool diagonal(int n, int base, const int *ai, const int *aj, const double *av,
              int *diagpos, double *diag);
void transpose(int rows, int cols, int base, const int *ai, const int *aj,
               const double *av, int *ti, int *tj, double *tv);
void permute(int n, int base, const int *ai, const int *aj, const double *av,
             const int *iperm, const int *perm, int *pai, int *paj,
             double *pav);
// RCM, bit-exact with the reference's SerialCM: with a team of host threads
// (nthreads <= 0: LAAMD_HOST_THREADS, else the hardware's, at most 16) for
// structurally symmetric patterns, serially otherwise
void rcm(int n, int base, const int *ai, const int *aj, int *iperm, int *perm,
         int nthreads = 0);
void rcm_serial(int n, int base, const int *ai, const int *aj, int *iperm, int *perm);
bool pattern_symmetric(int n, int base, const int *ai, const int *aj);
void iluk_symbolic(int n, int base, const int *ai, const int *aj, int level,
                   std::vector<int> &fai, std::vector<int> &faj,
                   std::vector<int> &first_upper);
bool iluk_numeric(int n, int base, const int *ai, const int *aj,
                  const double *av, const int *fai, const int *faj,
                  const int *first_upper, double *fav);
void icc_symbolic(int n, int base, const int *ai, const int *aj,
                  const int *diag_pos, int lvl, std::vector<int> &icc_ai,
                  std::vector<int> &icc_aj);
bool icc_numeric(int n, int base, const int *ai, const int *aj, const double *av,
                 const int *diag_pos, double omega, const int *icc_ai,
                 const int *icc_aj, double *icc_av);
bool ick_numeric(int n, int base, const int *ai, const int *aj, const double *av,
                 const int *diag_pos, double shift, const int *icc_ai,
                 const int *icc_aj, double *icc_av);

// Tiled schedule of the sync-free triangular solve (tiling.cpp): rows cut
// into tiles of a grid the pattern is recognised as, each tile's rows in
// (level, index) order in groups of at most `lanes` / rlanes rows of one
// level. Rows of up to kTileMaxEntries entries take one lane; longer rows,
// up to kTileRowLanes * kTileMaxEntries, take kTileRowLanes lanes.
constexpr int kTileMaxEntries = 8;  // entries per lane
constexpr int kTileRowLanes = 4;    // lanes of a long row
constexpr int kTileWindow = 16;     // groups a tile keeps its results for
struct TileLayout {
  int nparts = 0, ngroups = 0;
  int rlanes = 1;  // lanes per row (a row's entries are split in order)
  int m_max = 0;   // rows of the largest tile
  int h_max = 0;   // most values one tile reads from other tiles
  int maxe = 0;    // entries per lane slot (4 or kTileMaxEntries)
  int tau = 0;     // tile edge in grid steps
  double remote_frac = 0;  // entries read from another tile
  std::vector<int> part_gptr;  // tile p: groups [part_gptr[p], part_gptr[p+1])
  std::vector<int> grp_row;    // group g: row positions [grp_row[g], grp_row[g+1])
  std::vector<int> tperm;      // row position -> row
  // per row position, rlanes slots of maxe entries: >= 0 an LDS slot of
  // the tile's own earlier rows ((group % kTileWindow) * lanes + rlanes *
  // row in group; the slot kTileWindow * lanes reads 0 and pads short
  // rows), < 0 the halo slot -(code+1); tsrc: the entry's index in the
  // row-order triangle (-1: pad)
  std::vector<int> tcode, tsrc;
  // group g's first reads from other tiles: [rf_ptr[g], rf_ptr[g+1]) of
  // (rf_col: the row, rf_slot: its halo slot)
  std::vector<int> rf_ptr, rf_col, rf_slot;
};
// ai/aj: 0-based strict triangle in row order; level: each row's level.
// False (with the reason) when no tiling applies.
bool build_tile_layout(int n, const int *ai, const int *aj, const int *level, int lanes,
                       int max_parts, int max_rows, int max_halo, TileLayout &out,
                       std::string &why);

// seeded generators (gen.cpp)
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  // uniform in [0,1) with 53 random bits
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace laamd
