// Seeded synthetic inputs of the benchmark workloads: 2-D/3-D Poisson
// stencils, a power-law random matrix, a 27-point anisotropic
// variable-coefficient diffusion operator and a 2-D edge-conductivity
// diffusion operator. Every random draw is counter-based (SplitMix64 of
// seed + index), so results do not depend on thread count or call order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "host_algos.hpp"

namespace laamd {

namespace {
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ULL;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
// k-th output of the SplitMix64 stream started at `seed`
inline uint64_t stream_at(uint64_t seed, uint64_t k) { return mix64(seed + (k + 1) * kGamma); }
inline double u01(uint64_t bits) { return (bits >> 11) * (1.0 / 9007199254740992.0); }
// standard normal from two counter-based uniforms (Box-Muller, cosine branch)
inline double normal_at(uint64_t seed, uint64_t k) {
  double u1 = u01(stream_at(seed, 2 * k)), u2 = u01(stream_at(seed, 2 * k + 1));
  if (u1 < 1e-300) u1 = 1e-300;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
}  // namespace

void gen_uniform(int64_t n, uint64_t seed, double lo, double hi, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) out[k] = lo + (hi - lo) * u01(stream_at(seed, k));
}

void gen_lognormal(int64_t n, uint64_t seed, double mu, double sigma, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) out[k] = std::exp(mu + sigma * normal_at(seed, k));
}

// 2-D 5-point Dirichlet Poisson, natural order, stencil [4,-1,-1,-1,-1].
void gen_poisson2d(int nx, int ny, int *ai, int *aj, double *av, int64_t *nnz) {
  int64_t p = 0;
  ai[0] = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int r = i + nx * j;
      auto put = [&](int c, double v) {
        if (aj) { aj[p] = c; av[p] = v; }
        ++p;
      };
      if (j > 0) put(r - nx, -1.0);
      if (i > 0) put(r - 1, -1.0);
      put(r, 4.0);
      if (i + 1 < nx) put(r + 1, -1.0);
      if (j + 1 < ny) put(r + nx, -1.0);
      ai[r + 1] = static_cast<int>(p);
    }
  if (nnz) *nnz = p;
}

// 3-D 7-point Dirichlet Laplacian, natural order, diagonal 6.
void gen_laplace3d(int nx, int ny, int nz, int *ai, int *aj, double *av, int64_t *nnz) {
  const int64_t n = static_cast<int64_t>(nx) * ny * nz;
  // closed-form row starts so rows can be filled in parallel
  auto row_len = [&](int i, int j, int k) {
    return 1 + (i > 0) + (i + 1 < nx) + (j > 0) + (j + 1 < ny) + (k > 0) + (k + 1 < nz);
  };
  ai[0] = 0;
  {
    int64_t p = 0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          p += row_len(i, j, k);
          ai[i + static_cast<int64_t>(nx) * (j + static_cast<int64_t>(ny) * k) + 1] = static_cast<int>(p);
        }
  }
  if (nnz) *nnz = ai[n];
  if (!aj) return;
  const int64_t sxy = static_cast<int64_t>(nx) * ny;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t r = i + static_cast<int64_t>(nx) * j + sxy * k;
        int64_t p = ai[r];
        auto put = [&](int64_t c, double v) { aj[p] = static_cast<int>(c); av[p] = v; ++p; };
        if (k > 0) put(r - sxy, -1.0);
        if (j > 0) put(r - nx, -1.0);
        if (i > 0) put(r - 1, -1.0);
        put(r, 6.0);
        if (i + 1 < nx) put(r + 1, -1.0);
        if (j + 1 < ny) put(r + nx, -1.0);
        if (k + 1 < nz) put(r + sxy, -1.0);
      }
}

// Power-law rows: length ~ lognormal with the given mean (mu = ln(mean) -
// sigma^2/2), clipped to [1, max_len]; length-1 off-diagonal columns drawn
// uniformly and deduplicated (the diagonal removed if drawn), values N(0,1),
// then the diagonal set to 1 + sum |offdiag| (strict diagonal dominance).
// Each row draws from its own counter-based stream, so phase one (ai only)
// and phase two (aj/av) agree.
void gen_powerlaw(int n, double mean_len, double sigma, int max_len, uint64_t seed,
                  int *ai, int *aj, double *av, int64_t *nnz) {
  const double mu = std::log(mean_len) - 0.5 * sigma * sigma;
  const uint64_t len_seed = mix64(seed ^ 0x1111), col_seed = mix64(seed ^ 0x2222),
                 val_seed = mix64(seed ^ 0x3333);
  std::vector<int> lens(n);
  auto row_cols = [&](int r, std::vector<int> &cols) {
    double z = normal_at(len_seed, r);
    double L = std::round(std::exp(mu + sigma * z));
    int len = static_cast<int>(std::min<double>(std::max<double>(L, 1.0), max_len));
    cols.clear();
    const uint64_t rs = mix64(col_seed + static_cast<uint64_t>(r) * kGamma);
    for (int t = 0; t < len - 1; ++t) {
      int c = static_cast<int>(u01(stream_at(rs, t)) * n);
      if (c >= n) c = n - 1;
      if (c != r) cols.push_back(c);
    }
    cols.push_back(r);
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
  };
  if (!aj) {
#pragma omp parallel
    {
      std::vector<int> cols;
#pragma omp for schedule(dynamic, 4096)
      for (int r = 0; r < n; ++r) {
        row_cols(r, cols);
        lens[r] = static_cast<int>(cols.size());
      }
    }
    int64_t p = 0;
    ai[0] = 0;
    for (int r = 0; r < n; ++r) {
      p += lens[r];
      LA_CHECK_ARG(p < INT32_MAX, "power-law matrix exceeds int32 nnz");
      ai[r + 1] = static_cast<int>(p);
    }
    if (nnz) *nnz = p;
    return;
  }
#pragma omp parallel
  {
    std::vector<int> cols;
#pragma omp for schedule(dynamic, 4096)
    for (int r = 0; r < n; ++r) {
      row_cols(r, cols);
      int64_t p = ai[r];
      const uint64_t vs = mix64(val_seed + static_cast<uint64_t>(r) * kGamma);
      double off = 0;
      int64_t dpos = -1;
      for (size_t t = 0; t < cols.size(); ++t, ++p) {
        aj[p] = cols[t];
        if (cols[t] == r) {
          dpos = p;
          av[p] = 0;
        } else {
          av[p] = normal_at(vs, t);
          off += std::fabs(av[p]);
        }
      }
      av[dpos] = 1.0 + off;
    }
  }
  if (nnz) *nnz = ai[n];
}

// 27-point variable-coefficient anisotropic diffusion on an nx*ny*nz cell
// grid with Dirichlet ghosts: per-cell kappa ~ lognormal(0,1); the coupling
// of cells p,q at offset d in {-1,0,1}^3 is hm(kappa_p,kappa_q) * g(d) with
// hm the harmonic mean and g(d) = (dx^2 + dy^2 + aniso*dz^2) / |d|^4; a
// missing neighbour adds kappa_p * g(d) to the diagonal. The result is a
// symmetric, weakly diagonally dominant M-matrix (SPD).
void gen_aniso27(int nx, int ny, int nz, double aniso, uint64_t seed, int *ai, int *aj,
                 double *av, int64_t *nnz) {
  const int64_t sxy = static_cast<int64_t>(nx) * ny, n = sxy * nz;
  auto inside = [&](int i, int j, int k) {
    return i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz;
  };
  if (!aj) {
    ai[0] = 0;
    int64_t p = 0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          int cnt = 0;
          for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dx = -1; dx <= 1; ++dx) cnt += inside(i + dx, j + dy, k + dz);
          p += cnt;
          LA_CHECK_ARG(p < INT32_MAX, "27-point matrix exceeds int32 nnz");
          ai[i + nx * static_cast<int64_t>(j) + sxy * k + 1] = static_cast<int>(p);
        }
    if (nnz) *nnz = p;
    return;
  }
  std::vector<double> kappa(n);
  gen_lognormal(n, seed, 0.0, 1.0, kappa.data());
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t r = i + nx * static_cast<int64_t>(j) + sxy * k;
        int64_t p = ai[r], dpos = -1;
        double diag = 0;
        const double kp = kappa[r];
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              if (dx == 0 && dy == 0 && dz == 0) {
                dpos = p;
                aj[p] = static_cast<int>(r);
                av[p++] = 0;
                continue;
              }
              const double d2 = dx * dx + dy * dy + dz * dz;
              const double g = (dx * dx + dy * dy + aniso * dz * dz) / (d2 * d2);
              if (inside(i + dx, j + dy, k + dz)) {
                const int64_t q = r + dx + nx * static_cast<int64_t>(dy) + sxy * dz;
                const double kq = kappa[q];
                const double w = 2.0 * kp * kq / (kp + kq) * g;
                aj[p] = static_cast<int>(q);
                av[p++] = -w;
                diag += w;
              } else {
                diag += kp * g;
              }
            }
        av[dpos] = diag;
      }
  if (nnz) *nnz = ai[n];
}

// 2-D 5-point diffusion with per-edge conductivities theta. Edges: first
// the horizontal ones, e = j*(nx-1) + i joining (i,j)-(i+1,j), then the
// vertical ones, e = (nx-1)*ny + j*nx + i joining (i,j)-(i,j+1). Boundary
// nodes get a unit conductance to ground per missing neighbour.
void gen_diffusion2d(int nx, int ny, const double *theta, int *ai, int *aj, double *av,
                     int *edge_i, int *edge_j, int *val_map, int64_t *nnz) {
  const int n = nx * ny;
  const int64_t ne_h = static_cast<int64_t>(nx - 1) * ny;
  int64_t p = 0;
  ai[0] = 0;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int r = i + nx * j;
      const int nb[4] = {j > 0 ? r - nx : -1, i > 0 ? r - 1 : -1, i + 1 < nx ? r + 1 : -1,
                         j + 1 < ny ? r + nx : -1};
      auto edge_of = [&](int which) -> int64_t {  // edge id for neighbour slot
        switch (which) {
          case 0: return ne_h + static_cast<int64_t>(j - 1) * nx + i;  // south
          case 1: return static_cast<int64_t>(j) * (nx - 1) + (i - 1);  // west
          case 2: return static_cast<int64_t>(j) * (nx - 1) + i;        // east
          default: return ne_h + static_cast<int64_t>(j) * nx + i;      // north
        }
      };
      double diag = 0;
      int64_t dpos = -1;
      for (int s = 0; s < 5; ++s) {
        if (s == 2) {  // diagonal sits between west and east
          dpos = p;
          if (aj) aj[p] = r;
          ++p;
          continue;
        }
        const int slot = s < 2 ? s : s - 1;
        const int c = nb[slot];
        if (c < 0) {
          diag += 1.0;
          continue;
        }
        const int64_t e = edge_of(slot);
        if (aj) {
          aj[p] = c;
          av[p] = -theta[e];
          // (i,j) entry belongs to the edge's lower end if r < c
          if (r < c) { edge_i[e] = r; edge_j[e] = c; val_map[4 * e + 0] = static_cast<int>(p); }
          else val_map[4 * e + 1] = static_cast<int>(p);
        }
        diag += theta[e];
        ++p;
      }
      if (aj) av[dpos] = diag;
      ai[r + 1] = static_cast<int>(p);
    }
  if (aj) {
    // diagonal positions of both ends
    const int64_t ne = ne_h + static_cast<int64_t>(nx) * (ny - 1);
    for (int64_t e = 0; e < ne; ++e) {
      const int a = edge_i[e], b = edge_j[e];
      for (int q = ai[a]; q < ai[a + 1]; ++q)
        if (aj[q] == a) val_map[4 * e + 2] = q;
      for (int q = ai[b]; q < ai[b + 1]; ++q)
        if (aj[q] == b) val_map[4 * e + 3] = q;
    }
  }
  (void)n;
  if (nnz) *nnz = p;
}

}  // namespace laamd
