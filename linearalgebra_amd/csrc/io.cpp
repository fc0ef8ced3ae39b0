// Matrix readers: Matrix Market coordinate files and the reference's binary
// COO format, both into a CSR with rows sorted by (row, column).
//  * Matrix Market (utils/utils.h:93-132 with fast_matrix_market as the
//    triplet reader): coordinate real / integer / pattern (value 1),
//    general / symmetric / skew-symmetric; symmetric files are expanded to
//    both triangles (a diagonal entry is stored once), skew-symmetric ones
//    with the mirrored entry negated. Triplets are sorted by (row, column)
//    and counted into row pointers.
//  * Binary COO (utils/utils.h:24-91, ReadFromBinaryCOO / ReadFromBinaryCSR):
//    an int64 count, then that many 16-byte records laid out as
//    std::tuple<int32 row, int32 col, double val> is in libstdc++ (value at
//    byte 0, column at 8, row at 12), indices in the file's base. The matrix
//    is (max row + 1 - base) x (max col + 1 - base); indices keep the base.
//  * Binary vector (utils/utils.cpp:36-46): int64 count, then the doubles.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "host_algos.hpp"

namespace laamd {

namespace {
struct Triplets {
  std::vector<int> r, c;
  std::vector<double> v;
};

// sorted by (row, column), stable for equal keys; CSR with the given base
void to_csr(int rows, Triplets &t, int base, std::vector<int> &ai, std::vector<int> &aj,
            std::vector<double> &av, int index_base_in) {
  const size_t m = t.r.size();
  std::vector<size_t> idx(m);
  std::iota(idx.begin(), idx.end(), size_t(0));
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    return t.r[a] != t.r[b] ? t.r[a] < t.r[b] : t.c[a] < t.c[b];
  });
  ai.assign(static_cast<size_t>(rows) + 1, 0);
  aj.resize(m);
  av.resize(m);
  for (size_t k = 0; k < m; ++k) {
    const size_t s = idx[k];
    const int r0 = t.r[s] - index_base_in;
    LA_CHECK_ARG(r0 >= 0 && r0 < rows, "row index out of range");
    ++ai[static_cast<size_t>(r0) + 1];
    aj[k] = t.c[s] - index_base_in + base;
    av[k] = t.v[s];
  }
  for (int i = 0; i < rows; ++i) ai[i + 1] += ai[i];
  if (base)
    for (auto &p : ai) p += base;
}

std::string lower(std::string s) {
  for (auto &ch : s) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
  return s;
}
}  // namespace

void read_matrix_market(const char *path, int base, int *rows, int *cols, std::vector<int> &ai,
                        std::vector<int> &aj, std::vector<double> &av) {
  LA_CHECK_ARG(path && (base == 0 || base == 1), "bad argument");
  std::ifstream in(path);
  if (!in) throw la_error(LA_ERR_ARG, std::string("cannot open ") + path);
  std::string line;
  if (!std::getline(in, line)) throw la_error(LA_ERR_ARG, "empty Matrix Market file");
  std::istringstream hs(line);
  std::string banner, object, format, field, symmetry;
  hs >> banner >> object >> format >> field >> symmetry;
  LA_CHECK_ARG(lower(banner) == "%%matrixmarket" && lower(object) == "matrix",
               "not a Matrix Market matrix file");
  LA_CHECK_ARG(lower(format) == "coordinate", "only coordinate Matrix Market files are read");
  field = lower(field);
  symmetry = lower(symmetry);
  LA_CHECK_ARG(field == "real" || field == "integer" || field == "pattern" || field == "double",
               "Matrix Market field must be real, integer or pattern");
  LA_CHECK_ARG(symmetry == "general" || symmetry == "symmetric" || symmetry == "skew-symmetric",
               "Matrix Market symmetry must be general, symmetric or skew-symmetric");
  const bool pattern = field == "pattern";
  // size line after comments
  long long m = 0, n = 0, nz = 0;
  while (std::getline(in, line)) {
    size_t p = line.find_first_not_of(" \t\r");
    if (p == std::string::npos || line[p] == '%') continue;
    std::istringstream ss(line);
    ss >> m >> n >> nz;
    LA_CHECK_ARG(ss && m >= 0 && n >= 0 && nz >= 0 && m < (1LL << 31) && n < (1LL << 31),
                 "bad Matrix Market size line");
    break;
  }
  Triplets t;
  const size_t cap = static_cast<size_t>(symmetry == "general" ? nz : 2 * nz);
  t.r.reserve(cap);
  t.c.reserve(cap);
  t.v.reserve(cap);
  long long got = 0;
  while (got < nz && std::getline(in, line)) {
    size_t p = line.find_first_not_of(" \t\r");
    if (p == std::string::npos || line[p] == '%') continue;
    const char *s = line.c_str() + p;
    char *e = nullptr;
    const long long i = std::strtoll(s, &e, 10);
    const long long j = std::strtoll(e, &e, 10);
    double v = 1.0;
    if (!pattern) v = std::strtod(e, &e);
    LA_CHECK_ARG(i >= 1 && i <= m && j >= 1 && j <= n, "Matrix Market entry out of range");
    t.r.push_back(static_cast<int>(i - 1));
    t.c.push_back(static_cast<int>(j - 1));
    t.v.push_back(v);
    ++got;
  }
  LA_CHECK_ARG(got == nz, "Matrix Market file ends before its entries");
  if (symmetry != "general") {
    const double sgn = symmetry == "skew-symmetric" ? -1.0 : 1.0;
    const size_t k = t.r.size();
    for (size_t q = 0; q < k; ++q)
      if (t.r[q] != t.c[q]) {
        t.r.push_back(t.c[q]);
        t.c.push_back(t.r[q]);
        t.v.push_back(sgn * t.v[q]);
      }
  }
  *rows = static_cast<int>(m);
  *cols = static_cast<int>(n);
  to_csr(static_cast<int>(m), t, base, ai, aj, av, 0);
}

void read_binary_coo_csr(const char *path, int base, int *rows, int *cols, std::vector<int> &ai,
                         std::vector<int> &aj, std::vector<double> &av) {
  LA_CHECK_ARG(path && (base == 0 || base == 1), "bad argument");
  std::ifstream in(path, std::ios::binary);
  if (!in) throw la_error(LA_ERR_ARG, std::string("cannot open ") + path);
  int64_t size = 0;
  in.read(reinterpret_cast<char *>(&size), sizeof size);
  LA_CHECK_ARG(in && size >= 0, "bad binary COO header");
  Triplets t;
  t.r.resize(static_cast<size_t>(size));
  t.c.resize(static_cast<size_t>(size));
  t.v.resize(static_cast<size_t>(size));
  int mr = 0, mc = 0;
  for (int64_t q = 0; q < size; ++q) {
    unsigned char rec[16];
    in.read(reinterpret_cast<char *>(rec), sizeof rec);
    LA_CHECK_ARG(static_cast<bool>(in), "binary COO file ends before its records");
    double v;
    int32_t c, r;
    std::memcpy(&v, rec, 8);
    std::memcpy(&c, rec + 8, 4);
    std::memcpy(&r, rec + 12, 4);
    LA_CHECK_ARG(r >= base && c >= base, "binary COO index below the base");
    t.r[q] = r;
    t.c[q] = c;
    t.v[q] = v;
    mr = std::max(mr, r);
    mc = std::max(mc, c);
  }
  *rows = mr + (1 - base);
  *cols = mc + (1 - base);
  // entries keep the file's base: convert to 0-based rows for counting,
  // columns as stored
  to_csr(*rows, t, base, ai, aj, av, base);
}

void read_binary_vector(const char *path, std::vector<double> &v) {
  LA_CHECK_ARG(path, "bad argument");
  std::ifstream in(path, std::ios::binary);
  if (!in) throw la_error(LA_ERR_ARG, std::string("cannot open ") + path);
  int64_t size = 0;
  in.read(reinterpret_cast<char *>(&size), sizeof size);
  LA_CHECK_ARG(in && size >= 0, "bad binary vector header");
  v.resize(static_cast<size_t>(size));
  if (size) in.read(reinterpret_cast<char *>(v.data()), sizeof(double) * size);
  LA_CHECK_ARG(static_cast<bool>(in), "binary vector file ends early");
}

}  // namespace laamd
