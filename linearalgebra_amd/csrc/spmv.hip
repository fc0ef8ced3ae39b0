// CSR SpMV on gfx950.
//
// Kernel families (all compute y = alpha*A*x + beta*z):
//  * row kernel: G lanes of a 64-wide wavefront per row, G the power of two
//    at or above the mean row length; coalesced row streams, DPP reduction.
//    Replaces SerialSPMV/ParallelSPMV (sparse_mat_op/spmv.hpp:12-57).
//  * segment kernel: every wavefront owns a fixed slice of 64*E nonzeros
//    (nnz-balanced, like ALBUSSPMV's equal nnz split, spmv.hpp:195-228).
//    Products are staged in LDS; the rows that start inside the slice are
//    reduced from LDS, the partial head/tail rows leave one partial each and
//    a second tiny kernel adds them in slice order (the ALBUS boundary fix-up
//    of spmv.hpp:275-292, deterministic). Replaces SegSumSPMV and ALBUSSPMV.
//  * sliced ELL: one lane per row of 64-row column-major slices (near-uniform
//    row lengths).
//  * row panels x column blocks: a workgroup per panel of rows, its sums in
//    LDS across blocks of columns (ragged rows scattered over a large x).
#include <algorithm>
#include <climits>
#include <cstdlib>

#include "dev_util.hpp"
#include "device.hpp"
#include "host_algos.hpp"

namespace laamd {
using namespace dev;

namespace {
constexpr int kBlock = 256;
constexpr int kSegE = 8;               // nonzeros per lane per segment
constexpr int kSegNnz = 64 * kSegE;    // nonzeros per wavefront segment

template <typename V>
__device__ __forceinline__ V ldx(const V *x, int j) {
  return x[j];
}

// ------------------------------------------------------------------ row
// G lanes per row (lane sub adds entries sub, sub + G, ... in order, then a
// DPP sum over the G lanes); each lane group takes RB rows, RB * 64 / G
// consecutive rows per wavefront, and issues the first entry's loads of all
// of them before any product: with rows of a few entries a wavefront
// otherwise has too few bytes in flight to keep HBM busy.
template <int G, int RB, typename T, typename V, typename A>
__global__ __launch_bounds__(kBlock) void k_spmv_row(int n, const int *__restrict__ ai,
                                                     const int *__restrict__ aj,
                                                     const T *__restrict__ av,
                                                     const V *__restrict__ x, V *y,
                                                     const V *z, A alpha, A beta) {
  constexpr int NG = 64 / G;  // lane groups per wavefront
  const int lane = threadIdx.x & 63, g = lane / G, sub = lane & (G - 1);
  const int64_t base =
      (static_cast<int64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6)) * NG * RB + g;
  int rs[RB], re[RB], c[RB];
  A v[RB];
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const int64_t row = base + k * NG;
    rs[k] = re[k] = 0;
    if (row < n) {
      rs[k] = ai[row];
      re[k] = ai[row + 1];
    }
  }
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const int q = rs[k] + sub;
    c[k] = 0;
    v[k] = A(0);
    if (q < re[k]) {
      c[k] = ldi(aj, q);
      v[k] = A(ldm(av, q));
    }
  }
  A acc[RB];
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    acc[k] = A(0);
    if (rs[k] + sub < re[k]) acc[k] += v[k] * A(ldx(x, c[k]));
  }
#pragma unroll
  for (int k = 0; k < RB; ++k)
    for (int q = rs[k] + sub + G; q < re[k]; q += G) acc[k] += A(ldm(av, q)) * A(ldx(x, ldi(aj, q)));
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const int64_t row = base + k * NG;
    acc[k] = group_sum<G>(acc[k]);
    if (sub == 0 && row < n) {
      A r = alpha * acc[k];
      if (beta != A(0)) r += beta * A(z[row]);
      y[row] = V(r);
    }
  }
}

// ------------------------------------------------------------------ segment
// first_row[s] = first row r with ai[r] >= s*kSegNnz (r = n past the end);
// first_row[nseg] = n.
__global__ void k_seg_table(int n, int64_t nseg, const int *__restrict__ ai, int *first_row) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s > nseg) return;
  if (s == nseg) { first_row[s] = n; return; }
  const int64_t g = s * kSegNnz;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (static_cast<int64_t>(ai[mid]) < g) lo = mid + 1; else hi = mid;
  }
  first_row[s] = lo;
}

// Rows [rb, rb+R) owned by a segment, G lanes each, sums read from LDS. Row
// bounds of the first 128 owned rows come from registers preloaded before
// the product loop (pa = ai[f0+lane], pb = ai[f0+64+lane]); later chunks load
// them. Every lane executes the shuffles; groups past R drop out after.
template <int G, typename V, typename A>
__device__ __forceinline__ void seg_rows(int lane, int chunk, int rb, int R, int64_t g0,
                                         int64_t g1, int pa, int pb,
                                         const int *__restrict__ ai, const A *p, V *y,
                                         const V *z, A alpha, A beta, A *tail_part,
                                         int64_t s) {
  const int gi = lane / G, sub = lane & (G - 1);
  int rs = 0, re = 0;
  if (chunk == 0) {
    rs = __shfl(pa, gi, 64);
    const int nx = __shfl(pa, (gi + 1) & 63, 64), nb = __shfl(pb, 0, 64);
    re = gi + 1 < 64 ? nx : nb;
  } else if (chunk == 1) {
    rs = __shfl(pb, gi, 64);
    const int nx = __shfl(pb, (gi + 1) & 63, 64);
    re = gi + 1 < 64 ? nx : (gi < R ? ai[rb + 64] : 0);
  } else if (gi < R) {
    rs = ai[rb + gi];
    re = ai[rb + gi + 1];
  }
  if (gi >= R) return;  // whole groups only
  const int r = rb + gi;
  const int64_t qe = re < g1 ? re : g1;
  A acc = 0;
  for (int64_t q = rs + sub; q < qe; q += G) acc += p[q - g0];
  acc = group_sum<G>(acc);
  if (sub == 0) {
    if (re <= g1) {
      A o = alpha * acc;
      if (beta != A(0)) o += beta * A(z[r]);
      y[r] = V(o);
    } else {
      tail_part[s] = acc;  // continues into the next segment(s)
    }
  }
}

// four consecutive matrix entries / column indices per lane, one vector
// load each (non-temporal: streamed once per product)
template <typename T> struct Quad;
template <> struct Quad<double> { typedef double type __attribute__((ext_vector_type(4))); };
template <> struct Quad<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Quad<bf16> { typedef unsigned short type __attribute__((ext_vector_type(4))); };
typedef int int4v __attribute__((ext_vector_type(4)));

template <typename A>
__device__ __forceinline__ A widen(double v) { return A(v); }
template <typename A>
__device__ __forceinline__ A widen(float v) { return A(v); }
template <typename A>
__device__ __forceinline__ A widen(unsigned short v) {
  return A(__uint_as_float(static_cast<uint32_t>(v) << 16));
}

template <typename T, typename V, typename A>
__global__ __launch_bounds__(kBlock) void k_spmv_seg(
    int n, int64_t nnz, int64_t nseg, const int *__restrict__ ai, const int *__restrict__ aj,
    const T *__restrict__ av, const V *__restrict__ x, const int *__restrict__ first_row,
    A *head_part, A *tail_part, int *tail_row, int *tail_end, V *y, const V *z, A alpha,
    A beta) {
  using Q = typename Quad<T>::type;
  static_assert(kSegE == 8, "two quads per lane");
  __shared__ A sp[kBlock / 64][kSegNnz];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t s = static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * (kBlock / 64) + wave;
  const bool live = s < nseg;
  const int64_t g0 = s * kSegNnz;
  const int64_t g1 = live ? (g0 + kSegNnz < nnz ? g0 + kSegNnz : nnz) : g0;
  A *p = sp[wave];
  // lane owns nonzeros g0 + h*256 + 4*lane + e, h in {0,1}, e in {0..3}.
  // Column and value arrays are padded with zeros to a whole number of
  // segments, so every load is an unconditional vector load (no branch, no
  // wait between them); products past the segment's end are zeroed. Waves
  // past the last segment load segment 0 and discard it.
  const int64_t gl = live ? g0 : 0;
  int c[8];
  Q v[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t g = gl + h * 256 + 4 * lane;
    const int4v cc = __builtin_nontemporal_load(reinterpret_cast<const int4v *>(aj + g));
    v[h] = __builtin_nontemporal_load(reinterpret_cast<const Q *>(av + g));
    c[4 * h + 0] = cc.x; c[4 * h + 1] = cc.y; c[4 * h + 2] = cc.z; c[4 * h + 3] = cc.w;
  }
  // segment bounds and the first 128 owned row pointers (clamped to the
  // owned range) go out while the matrix stream is in flight
  const int64_t sl = live ? s : 0;
  const int f0 = first_row[sl], f1 = first_row[sl + 1];
  V xv[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) xv[e] = ldx(x, c[e]);
  const int pa = ai[f0 + lane <= f1 ? f0 + lane : f1];
  const int pb = ai[f0 + 64 + lane <= f1 ? f0 + 64 + lane : f1];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t g = g0 + h * 256 + 4 * lane + e;
      const A prod = widen<A>(v[h][e]) * A(xv[4 * h + e]);
      p[h * 256 + 4 * lane + e] = g < g1 ? prod : A(0);
    }
  __syncthreads();
  if (!live) return;
  // head: the row that started in an earlier segment (ALBUS bv[2t])
  const int64_t hb = __shfl(pa, 0, 64);  // ai[f0]
  if (f0 > 0 && hb > g0) {
    const int64_t he = hb < g1 ? hb : g1;
    A acc = 0;
    for (int64_t q = g0 + lane; q < he; q += 64) acc += p[q - g0];
    acc = wave_sum(acc);
    if (lane == 0) head_part[s] = acc;
  }
  if (lane == 0) {
    const int e = ai[f1];
    const bool crosses = f1 > f0 && static_cast<int64_t>(e) > g1;
    tail_row[s] = crosses ? f1 - 1 : -1;
    tail_end[s] = e;
  }
  int chunk = 0;
  for (int rb = f0; rb < f1; rb += 64, ++chunk) {
    const int R = (f1 - rb) < 64 ? (f1 - rb) : 64;
    const int Gs = R > 32 ? 1 : R > 16 ? 2 : R > 8 ? 4 : R > 4 ? 8 : R > 2 ? 16 : R > 1 ? 32 : 64;
#define LA_SEGROWS(GG) \
  seg_rows<GG>(lane, chunk, rb, R, g0, g1, pa, pb, ai, p, y, z, alpha, beta, tail_part, s)
    switch (Gs) {
      case 1: LA_SEGROWS(1); break;
      case 2: LA_SEGROWS(2); break;
      case 4: LA_SEGROWS(4); break;
      case 8: LA_SEGROWS(8); break;
      case 16: LA_SEGROWS(16); break;
      case 32: LA_SEGROWS(32); break;
      default: LA_SEGROWS(64); break;
    }
#undef LA_SEGROWS
  }
}

// rows that cross segment boundaries: tail partial + the following head
// partials, added in segment order. The row's end comes from the segment
// kernel, so every load but the rare third head goes out at once.
template <typename V, typename A>
__global__ void k_spmv_seg_fix(int64_t nseg, const int *__restrict__ tail_row,
                               const int *__restrict__ tail_end, const A *head_part,
                               const A *tail_part, V *y, const V *z, A alpha, A beta) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const int r = tail_row[s];
  const int64_t end = tail_end[s];
  const A tp = tail_part[s];
  const A hp = s + 1 < nseg ? head_part[s + 1] : A(0);
  if (r < 0) return;
  A acc = tp + hp;
  for (int64_t t = s + 2; t < nseg && end > t * kSegNnz; ++t) acc += head_part[t];
  A o = alpha * acc;
  if (beta != A(0)) o += beta * A(z[r]);
  y[r] = V(o);
}

// --------------------------------------------- row panels x column blocks
// For products whose x gathers scatter over an x far larger than an XCD's
// 4 MB L2. Columns are cut into blocks of at most 2^kPanelColBits columns
// (4 MB of FP64 x) and rows into panels of at most kPanelRows rows. One
// workgroup owns a panel: its row sums live in LDS while it sweeps the
// blocks in column order, so every workgroup gathers from the same block at
// about the same time (that block's lines stay in the L2s) and y is written
// once. Inside a (panel, block) tile the entries are in CSR order and cut
// at row boundaries into one chunk per wavefront (padded to a multiple of
// four with kPanelPad entries), so a row of a tile belongs to one
// wavefront. An entry is one 32-bit word: the row inside the panel above
// kPanelColBits bits of column inside the block.
constexpr int kPanelWaves = 8;
constexpr int kPanelThreads = 64 * kPanelWaves;
constexpr int kPanelRows = 8192;  // FP64 row sums: 64 KB of LDS per workgroup
constexpr int kPanelColBits = 19;
constexpr unsigned kPanelPad = 0xFFFFFFFFu;

template <int CTRL, int RM>
__device__ __forceinline__ int dpp_rm(int v, int old) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, RM, 0xF, false);
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_rm(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, RM, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, RM, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
template <int CTRL, int RM>
__device__ __forceinline__ float dpp_rm(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, RM, 0xF, false));
}
template <int CTRL, int RM, typename A>
__device__ __forceinline__ void seg_scan_step(A &v, int key) {
  const A t = dpp_rm<CTRL, RM>(v);
  const int k = dpp_rm<CTRL, RM>(key, INT_MIN);
  if (k == key) v += t;
}
// Inclusive scan of v over the lanes of a wavefront, restarted wherever the
// key changes (equal keys occupy consecutive lanes): row_shr 1, 2, 4, 8
// inside each 16-lane row, then row_bcast 15 (rows 1 and 3) and row_bcast 31
// (rows 2 and 3). Every lane must be active.
template <typename A>
__device__ __forceinline__ A seg_scan(A v, int key) {
  seg_scan_step<0x111, 0xF>(v, key);
  seg_scan_step<0x112, 0xF>(v, key);
  seg_scan_step<0x114, 0xF>(v, key);
  seg_scan_step<0x118, 0xF>(v, key);
  seg_scan_step<0x142, 0xA>(v, key);
  seg_scan_step<0x143, 0xC>(v, key);
  return v;
}

// Each lane takes four consecutive entries of its chunk (one vector load of
// words, one of values), sums the runs of equal rows in order, and adds the
// runs that begin and end inside its four straight into the row sums. A
// run that touches the lane's first or last entry may continue in the
// neighbouring lanes: the lanes' last runs are combined by a scan over
// lanes keyed by row, and each such row is added once, by the lane where
// it ends. Every row is therefore added to by one lane of one instruction,
// in an order fixed by the layout: the result is deterministic.
template <typename T, typename V, typename A>
__global__ __launch_bounds__(kPanelThreads) void k_spmv_panel(
    int nb, int width, const int *__restrict__ prow, const int *__restrict__ chunk,
    const unsigned *__restrict__ pidx, const T *__restrict__ pval, const V *__restrict__ x,
    V *y, const V *z, A alpha, A beta) {
  using Q = typename Quad<T>::type;
  __shared__ A ys[kPanelRows];
  const int p = blockIdx.x;
  const int r0 = prow[p], nr = prow[p + 1] - r0;
  for (int i = threadIdx.x; i < nr; i += kPanelThreads) ys[i] = A(0);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int *cp = chunk + static_cast<int64_t>(p) * nb * (kPanelWaves + 1) + wave;
  constexpr unsigned kColMask = (1u << kPanelColBits) - 1u;
  for (int b = 0; b < nb; ++b, cp += kPanelWaves + 1) {
    const int s = cp[0], e = cp[1];
    const V *xb = x + static_cast<int64_t>(b) * width;
    for (int q0 = s; q0 < e; q0 += 256) {
      const int q = q0 + 4 * lane;
      int rw[4];
      A pr[4];
      if (q < e) {  // chunks are whole quads
        const int4v c = __builtin_nontemporal_load(reinterpret_cast<const int4v *>(pidx + q));
        const Q v = __builtin_nontemporal_load(reinterpret_cast<const Q *>(pval + q));
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const unsigned u = static_cast<unsigned>(c[j]);
          const bool pad = u == kPanelPad;
          rw[j] = pad ? -1 : static_cast<int>(u >> kPanelColBits);
          const V xv = ldx(xb, pad ? 0 : static_cast<int>(u & kColMask));
          const A t = widen<A>(v[j]) * A(xv);
          pr[j] = pad ? A(0) : t;
        }
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) { rw[j] = -1; pr[j] = A(0); }
      }
      // runs of equal rows inside the lane; the first is held back (it may
      // continue an earlier lane's run), inner ones are complete
      A run = pr[0], hs = A(0);
      int cur = rw[0], hr = -1;
      bool multi = false;
#pragma unroll
      for (int j = 1; j < 4; ++j) {
        if (rw[j] == cur) {
          run += pr[j];
        } else {
          if (!multi) { hs = run; hr = cur; multi = true; }
          else if (cur >= 0) ys[cur] += run;
          cur = rw[j];
          run = pr[j];
        }
      }
      // the last run: scan over the lanes whose last runs share its row
      const A sc = seg_scan(run, cur);
      const A sp = __shfl_up(sc, 1, 64);
      const int kp = __shfl_up(cur, 1, 64);
      const int first = multi ? hr : cur;
      const int nfirst = __shfl_down(first, 1, 64);
      if (multi && hr >= 0) ys[hr] += (lane > 0 && kp == hr) ? sp + hs : hs;
      if (cur >= 0 && !(lane < 63 && nfirst == cur)) ys[cur] += sc;
    }
    __syncthreads();  // the next block's rows may fall to other wavefronts
  }
  for (int i = threadIdx.x; i < nr; i += kPanelThreads) {
    A o = alpha * ys[i];
    if (beta != A(0)) o += beta * A(z[r0 + i]);
    y[r0 + i] = V(o);
  }
}

// ------------------------------------------------------------------ grad
__global__ void k_value_grad(int n, const int *__restrict__ ai, const int *__restrict__ aj,
                             const double *__restrict__ lam, const double *__restrict__ x,
                             double *g) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double l = -lam[r];
  for (int q = ai[r]; q < ai[r + 1]; ++q) g[q] = l * x[aj[q]];
}

template <typename T, typename V, typename A>
void launch_row(Mat &M, const T *av, const V *x, V *y, const V *z, A alpha, A beta,
                hipStream_t s) {
  const int n = M.rows;
  int G = 1;
  while (G < 64 && G < M.avg_row) G <<= 1;
  const int *ai = M.ai.as<int>(), *aj = M.aj.as<int>();
  // rows per lane group, measured (FP64): 128^3 7-point (G 8) 72.8 us with
  // 1, 52.2 with 4, 66.1 with 8; 160^3 27-point (G 32) 538 us with 1, 453
  // with 2; the 4M power-law matrix (G 16) 1099 us with 1, 1116 with 2
#define LA_ROW(GG)                                                                           \
  {                                                                                          \
    constexpr int RB = GG <= 8 ? 4 : GG == 32 ? 2 : 1;                                       \
    const int64_t per = static_cast<int64_t>(kBlock / GG) * RB;                              \
    const unsigned grid = static_cast<unsigned>((n + per - 1) / per);                        \
    if (grid)                                                                                \
      hipLaunchKernelGGL((k_spmv_row<GG, RB, T, V, A>), dim3(grid), dim3(kBlock), 0, s, n, ai, \
                         aj, av, x, y, z, alpha, beta);                                      \
  }
  switch (G) {
    case 1: LA_ROW(1); break;
    case 2: LA_ROW(2); break;
    case 4: LA_ROW(4); break;
    case 8: LA_ROW(8); break;
    case 16: LA_ROW(16); break;
    case 32: LA_ROW(32); break;
    default: LA_ROW(64); break;
  }
#undef LA_ROW
  LA_HIP(hipGetLastError());
}

template <typename T, typename V, typename A>
void launch_seg(Mat &M, const T *av, const V *x, V *y, const V *z, A alpha, A beta,
                hipStream_t s) {
  if (!M.seg_ready) spmv_prepare_segments(M);
  const int64_t nseg = M.nseg;
  const unsigned grid = static_cast<unsigned>((nseg + 3) / 4);
  A *head = static_cast<A *>(M.seg_head.p), *tail = static_cast<A *>(M.seg_tail.p);
  hipLaunchKernelGGL((k_spmv_seg<T, V, A>), dim3(grid), dim3(kBlock), 0, s, M.rows, M.nnz,
                     nseg, M.ai.as<int>(), M.aj.as<int>(), av, x, M.seg_first_row.as<int>(),
                     head, tail, M.seg_tail_row.as<int>(), M.seg_tail_end.as<int>(), y, z,
                     alpha, beta);
  LA_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_spmv_seg_fix<V, A>), dim3((nseg + 255) / 256), dim3(256), 0, s, nseg,
                     M.seg_tail_row.as<int>(), M.seg_tail_end.as<int>(), head, tail, y, z,
                     alpha, beta);
  LA_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ SELL
// Sliced ELL: slices of 64 consecutive rows, each stored column-major with
// the slice's longest row as its width, so lane l of a wavefront walks row
// 64s+l left to right (SerialSPMV's order) with every load coalesced across
// the wavefront. Padding entries carry column -1 and add nothing (a
// non-finite x entry elsewhere cannot leak in); the copy's values are
// gathered from the CSR values of the same precision.
constexpr double kSellMaxFill = 1.15;  // padded entries / nnz allowed
constexpr int kSellMaxWidth = 64;      // longest row allowed
// Banded matrices (every |column - row| <= 32767, e.g. an RCM-ordered grid
// operator) store each slot's column as a 16-bit offset from its row: 2
// bytes per entry instead of 4 (a sixth of an FP64 SpMV's bytes, a third of
// a BF16 one's). Padding is the offset -32768.
constexpr int kSellPad16 = -32768;

// a slot's column: int32 columns as stored (padding -1), 16-bit offsets
// relative to the row
__device__ __forceinline__ int sell_colof(const int *__restrict__ c, int64_t q, int, bool nt) {
  return nt ? __builtin_nontemporal_load(c + q) : c[q];
}
__device__ __forceinline__ int sell_colof(const int16_t *__restrict__ c, int64_t q, int row,
                                          bool nt) {
  const int d = nt ? __builtin_nontemporal_load(c + q) : c[q];
  return d == kSellPad16 ? -1 : row + d;
}

// W consecutive slice columns of one lane: all loads issued before any
// product (the slice width is uniform across the wavefront)
// (NTC: the column indices with non-temporal loads too; measured on MI355X,
// narrow slices -- width <= 8, the 7-point stencil -- run faster with
// ordinary column loads, wide ones -- the 27-point stencil -- with
// non-temporal ones)
template <int W, bool NTC, typename I, typename T, typename V, typename A>
__device__ __forceinline__ A sell_block(A acc, int k0, int64_t base, int row,
                                        const I *__restrict__ scol,
                                        const T *__restrict__ sval, const V *__restrict__ x) {
  using LT = decltype(ldm(sval, int64_t(0)));  // bf16 loads widen to float
  int c[W];
  LT v[W];
#pragma unroll
  for (int e = 0; e < W; ++e) {
    const int64_t q = base + static_cast<int64_t>(k0 + e) * 64;
    c[e] = sell_colof(scol, q, row, NTC);
    v[e] = ldm(sval, q);
  }
  V xv[W];
#pragma unroll
  for (int e = 0; e < W; ++e) xv[e] = ldx(x, c[e] < 0 ? 0 : c[e]);
#pragma unroll
  for (int e = 0; e < W; ++e) {
    const A p = A(v[e]) * A(xv[e]);
    acc = c[e] >= 0 ? acc + p : acc;  // padding (column -1) contributes nothing
  }
  return acc;
}

template <typename I, typename T, typename V, typename A>
__global__ __launch_bounds__(kBlock) void k_spmv_sell(
    int n, int64_t nslices, int uw, const int64_t *__restrict__ sp,
    const int *__restrict__ sw, const I *__restrict__ scol, const T *__restrict__ sval,
    const V *__restrict__ x, V *y, const V *z, A alpha, A beta, const T *__restrict__ rs) {
  const int64_t sl = static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * (kBlock / 64) +
                     (threadIdx.x >> 6);
  if (sl >= nslices) return;
  const int lane = threadIdx.x & 63;
  const int row = static_cast<int>(sl * 64 + lane);
  // uw > 0: every slice has that width (no slice table to read)
  const int64_t base = (uw > 0 ? sl * 64 * uw : sp[sl]) + lane;
  const int w = uw > 0 ? uw : sw[sl];
  // the epilogue's operands are loaded first, so their latency overlaps the
  // row's loads instead of following them (z may alias y: each lane reads
  // its own element before writing it)
  A zr = 0, sr = 1;
  if (row < n) {
    if (beta != A(0)) zr = A(z[row]);
    if (rs) sr = static_cast<A>(ldm(rs, row));
  }
  A acc = 0;
#define LA_SB(WW, NT) acc = sell_block<WW, NT, I, T, V, A>(acc, k, base, row, scol, sval, x)
  int k = 0;
  if (w <= 8) {
    switch (w) {  // the whole row in one go
      case 8: LA_SB(8, false); break;
      case 7: LA_SB(7, false); break;
      case 6: LA_SB(6, false); break;
      case 5: LA_SB(5, false); break;
      case 4: LA_SB(4, false); break;
      case 3: LA_SB(3, false); break;
      case 2: LA_SB(2, false); break;
      case 1: LA_SB(1, false); break;
      default: break;
    }
  } else {
    for (; k + 16 <= w; k += 16) LA_SB(16, true);
    if (k + 8 <= w) {
      LA_SB(8, true);
      k += 8;
    }
    switch (w - k) {  // the rest: 0..7 columns
      case 7: LA_SB(7, true); break;
      case 6: LA_SB(6, true); break;
      case 5: LA_SB(5, true); break;
      case 4: LA_SB(4, true); break;
      case 3: LA_SB(3, true); break;
      case 2: LA_SB(2, true); break;
      case 1: LA_SB(1, true); break;
      default: break;
    }
  }
#undef LA_SB
  if (row < n) {
    A o = alpha * acc;
    if (beta != A(0)) o += beta * zr;
    if (rs) o *= sr;  // y = rs o (alpha A x + beta z)
    y[row] = V(o);
  }
}

template <typename T>
__global__ void k_sell_gather(int64_t m, const int *__restrict__ src, const T *__restrict__ v,
                              T *out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < m) {
    const int q = src[t];
    out[t] = q >= 0 ? v[q] : T();
  }
}

// builds the slice structure once; false when padding would cost too much
bool prepare_sell(Mat &M) {
  if (M.sell_state) return M.sell_state > 0;
  const int n = M.rows;
  const int64_t ns = (n + 63) / 64;
  std::vector<int64_t> sp(static_cast<size_t>(ns) + 1, 0);
  std::vector<int> sw(static_cast<size_t>(std::max<int64_t>(ns, 1)), 0);
  for (int64_t s = 0; s < ns; ++s) {
    int w = 0;
    for (int r = static_cast<int>(s * 64); r < std::min<int64_t>(n, (s + 1) * 64); ++r)
      w = std::max(w, M.h_ai[r + 1] - M.h_ai[r]);
    sw[s] = w;
    sp[s + 1] = sp[s] + 64 * static_cast<int64_t>(w);
  }
  const int64_t len = sp[ns];
  const int maxw = ns ? *std::max_element(sw.begin(), sw.begin() + ns) : 0;
  if (M.nnz == 0 || static_cast<double>(len) > kSellMaxFill * static_cast<double>(M.nnz) ||
      maxw > kSellMaxWidth || len >= INT32_MAX) {
    M.sell_state = -1;
    return false;
  }
  std::vector<int> col(static_cast<size_t>(len), -1), src(static_cast<size_t>(len), -1);
  for (int64_t s = 0; s < ns; ++s)
    for (int l = 0; l < 64; ++l) {
      const int64_t r = s * 64 + l;
      if (r >= n) break;
      const int b = M.h_ai[r], rl = M.h_ai[r + 1] - b;
      for (int k = 0; k < rl; ++k) {
        const int64_t p = sp[s] + static_cast<int64_t>(k) * 64 + l;
        col[p] = M.h_aj[b + k];
        src[p] = b + k;
      }
    }
  M.sell_nslices = ns;
  M.sell_len = len;
  M.sell_uniform_w = (ns && std::all_of(sw.begin(), sw.begin() + ns, [&](int w) { return w == sw[0]; }))
                         ? sw[0] : 0;
  // 16-bit offsets when every column is within 32767 of its row
  bool fits16 = true;
  for (int64_t s = 0; s < ns && fits16; ++s)
    for (int l = 0; l < 64; ++l) {
      const int64_t r = s * 64 + l;
      if (r >= n) break;
      for (int k = 0; k < sw[s]; ++k) {
        const int c = col[sp[s] + static_cast<int64_t>(k) * 64 + l];
        if (c >= 0 && std::abs(static_cast<int64_t>(c) - r) > 32767) {
          fits16 = false;
          break;
        }
      }
    }
  if (const char *e = std::getenv("LAAMD_SELL16"))
    if (e[0] == '0') fits16 = false;
  M.sell_col16 = fits16;
  M.sell_ptr.alloc(sizeof(int64_t) * sp.size());
  M.sell_ptr.upload(sp.data(), sizeof(int64_t) * sp.size(), M.stream);
  M.sell_w.alloc(sizeof(int) * sw.size());
  M.sell_w.upload(sw.data(), sizeof(int) * sw.size(), M.stream);
  if (fits16) {
    std::vector<int16_t> c16(col.size());
    for (int64_t s = 0; s < ns; ++s)
      for (int64_t p = sp[s]; p < sp[s + 1]; ++p) {
        const int64_t r = s * 64 + (p - sp[s]) % 64;
        c16[p] = col[p] < 0 ? static_cast<int16_t>(kSellPad16)
                            : static_cast<int16_t>(static_cast<int64_t>(col[p]) - r);
      }
    M.sell_col.alloc(sizeof(int16_t) * c16.size());
    M.sell_col.upload(c16.data(), sizeof(int16_t) * c16.size(), M.stream);
  } else {
    M.sell_col.alloc(sizeof(int) * col.size());
    M.sell_col.upload(col.data(), sizeof(int) * col.size(), M.stream);
  }
  M.sell_src.alloc(sizeof(int) * src.size());
  M.sell_src.upload(src.data(), sizeof(int) * src.size(), M.stream);
  LA_HIP(hipStreamSynchronize(M.stream));
  M.sell_state = 1;
  return true;
}

template <typename T>
const T *sell_values(Mat &M, la_dtype dt, const T *av, hipStream_t s) {
  if (M.sell_ver[dt] != M.vals.version || !M.sell_val[dt].p) {
    if (!M.sell_val[dt].p) M.sell_val[dt].alloc(sizeof(T) * static_cast<size_t>(M.sell_len));
    hipLaunchKernelGGL(k_sell_gather<T>, dim3(static_cast<unsigned>((M.sell_len + 255) / 256)),
                       dim3(256), 0, s, M.sell_len, M.sell_src.as<int>(), av,
                       M.sell_val[dt].as<T>());
    LA_HIP(hipGetLastError());
    M.sell_ver[dt] = M.vals.version;
  }
  return M.sell_val[dt].as<T>();
}

template <typename T> struct DtOf;
template <> struct DtOf<double> { static constexpr la_dtype v = LA_F64; };
template <> struct DtOf<float> { static constexpr la_dtype v = LA_F32; };
template <> struct DtOf<bf16> { static constexpr la_dtype v = LA_BF16; };

template <typename T, typename V, typename A>
void launch_sell(Mat &M, const T *av, const V *x, V *y, const V *z, A alpha, A beta,
                 hipStream_t s, const T *rs = nullptr) {
  const T *sv = sell_values<T>(M, DtOf<T>::v, av, s);
  constexpr int wpb = kBlock / 64;  // wavefronts per workgroup
  const unsigned grid = static_cast<unsigned>((M.sell_nslices + wpb - 1) / wpb);
  if (M.sell_col16)
    hipLaunchKernelGGL((k_spmv_sell<int16_t, T, V, A>), dim3(grid), dim3(kBlock), 0, s, M.rows,
                       M.sell_nslices, M.sell_uniform_w, M.sell_ptr.as<int64_t>(),
                       M.sell_w.as<int>(), M.sell_col.as<int16_t>(), sv, x, y, z, alpha, beta,
                       rs);
  else
    hipLaunchKernelGGL((k_spmv_sell<int, T, V, A>), dim3(grid), dim3(kBlock), 0, s, M.rows,
                       M.sell_nslices, M.sell_uniform_w, M.sell_ptr.as<int64_t>(),
                       M.sell_w.as<int>(), M.sell_col.as<int>(), sv, x, y, z, alpha, beta, rs);
  LA_HIP(hipGetLastError());
}

template <typename T, typename V, typename A>
void launch_panel(Mat &M, const void *pv, const void *x, void *y, const void *z, double alpha,
                  double beta, hipStream_t s) {
  hipLaunchKernelGGL((k_spmv_panel<T, V, A>), dim3(M.pb_np), dim3(kPanelThreads), 0, s,
                     M.pb_nb, M.pb_width, M.pb_prow.as<int>(), M.pb_chunk.as<int>(),
                     M.pb_idx.as<unsigned>(), static_cast<const T *>(pv),
                     static_cast<const V *>(x), static_cast<V *>(y), static_cast<const V *>(z),
                     A(alpha), A(beta));
  LA_HIP(hipGetLastError());
}

template <typename T, typename V, typename A>
void dispatch_alg(Mat &M, la_spmv_alg alg, const void *av, const void *x, void *y,
                  const void *z, double alpha, double beta, hipStream_t s) {
  if (M.nnz == 0 || M.cols == 0) alg = LA_SPMV_ROW;  // nothing to gather
  if (alg == LA_SPMV_AUTO) alg = prepare_sell(M) ? LA_SPMV_SELL : LA_SPMV_ALBUS;
  if (alg == LA_SPMV_SELL && !prepare_sell(M)) alg = LA_SPMV_ALBUS;
  const T *a = static_cast<const T *>(av);
  const V *xx = static_cast<const V *>(x);
  V *yy = static_cast<V *>(y);
  const V *zz = static_cast<const V *>(z);
  if (alg == LA_SPMV_ROW)
    launch_row<T, V, A>(M, a, xx, yy, zz, A(alpha), A(beta), s);
  else if (alg == LA_SPMV_SELL)
    launch_sell<T, V, A>(M, a, xx, yy, zz, A(alpha), A(beta), s);
  else
    launch_seg<T, V, A>(M, a, xx, yy, zz, A(alpha), A(beta), s);
}
}  // namespace

std::unique_ptr<Mat> make_mat(int rows, int cols, int base, const int *ai, const int *aj,
                              const double *av, hipStream_t s) {
  LA_CHECK_ARG(rows >= 0 && cols >= 0, "negative dimension");
  LA_CHECK_ARG(ai != nullptr, "ai is null");
  auto M = std::make_unique<Mat>();
  M->rows = rows;
  M->cols = cols;
  M->stream = s;
  const int64_t nnz = static_cast<int64_t>(ai[rows]) - ai[0];
  LA_CHECK_ARG(ai[0] == base, "ai[0] must equal base");
  LA_CHECK_ARG(nnz >= 0 && nnz < INT32_MAX, "nnz out of int32 range");
  LA_CHECK_ARG(nnz == 0 || (aj && av), "aj/av null");
  M->nnz = nnz;
  M->h_ai.resize(static_cast<size_t>(rows) + 1);
  M->h_aj.resize(static_cast<size_t>(nnz));
  int max_row = 0;
  for (int i = 0; i <= rows; ++i) M->h_ai[i] = ai[i] - base;
  for (int i = 0; i < rows; ++i) {
    const int len = M->h_ai[i + 1] - M->h_ai[i];
    LA_CHECK_ARG(len >= 0, "row pointers must be non-decreasing");
    max_row = std::max(max_row, len);
  }
  for (int64_t k = 0; k < nnz; ++k) {
    const int c = aj[k] - base;
    LA_CHECK_ARG(c >= 0 && c < cols, "column index out of range");
    M->h_aj[k] = c;
  }
  M->max_row = max_row;
  M->avg_row = rows ? static_cast<double>(nnz) / rows : 0.0;
  M->ai.alloc(sizeof(int) * (static_cast<size_t>(rows) + 1));
  const int64_t pnnz = padded_count(nnz);
  M->aj.alloc(sizeof(int) * static_cast<size_t>(pnnz));
  M->ai.upload(M->h_ai.data(), sizeof(int) * (rows + 1), s);
  M->aj.upload(M->h_aj.data(), sizeof(int) * nnz, s);
  if (pnnz > nnz)
    LA_HIP(hipMemsetAsync(M->aj.as<int>() + nnz, 0, sizeof(int) * (pnnz - nnz), s));
  M->vals.set_from_host(av, nnz, s);
  LA_HIP(hipStreamSynchronize(s));
  return M;
}

bool spmv_prepare_sell(Mat &A) { return prepare_sell(A); }

bool spmv_sell_rowscaled(Mat &A, la_dtype val, la_dtype vec, const void *rs, const void *x,
                         const void *z, void *y, hipStream_t s) {
  LA_CHECK_ARG(val >= LA_F64 && val <= LA_BF16 && (vec == LA_F64 || vec == LA_F32),
               "bad dtype");
  if (A.rows == 0 || A.nnz == 0 || A.cols == 0 || !prepare_sell(A)) return false;
  if (!A.vals.has[val]) A.vals.add(val, s);
  const void *av = A.vals.get(val);
#define LA_RS(T, V, AC)                                                                        \
  launch_sell<T, V, AC>(A, static_cast<const T *>(av), static_cast<const V *>(x),              \
                        static_cast<V *>(y), static_cast<const V *>(z), AC(-1), AC(1), s,      \
                        static_cast<const T *>(rs))
  if (vec == LA_F64) {
    if (val == LA_F64) LA_RS(double, double, double);
    else if (val == LA_F32) LA_RS(float, double, float);
    else LA_RS(bf16, double, float);
  } else {
    if (val == LA_F64) LA_RS(double, float, float);
    else if (val == LA_F32) LA_RS(float, float, float);
    else LA_RS(bf16, float, float);
  }
#undef LA_RS
  return true;
}

namespace {
// out[t] = v[src[t]] (src < 0: 0)
__global__ void k_gather_vals(int64_t m, const int *__restrict__ src,
                              const double *__restrict__ v, double *out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < m) {
    const int q = src[t];
    out[t] = q >= 0 ? v[q] : 0.0;
  }
}

// Column blocks. A product whose x gathers land all over an x much larger
// than an XCD's L2 (4 MB) fetches a 128-byte line for nearly every 4- or
// 8-byte value. The row-panel layout cuts the columns into blocks of
// kColBlockCols (512 KB of FP64 x) and keeps each panel's row sums in LDS
// across the blocks, so a block costs no pass over y. Measured on the
// 4M-row power-law matrix (FP64): blocks of 4 MB of x 540 us, 2 MB 435-439
// us, 1 MB 434 us, 768 KB 442 us, 512 KB 358-366 us, 384 KB 360 us, 256 KB
// 395 us, 128 KB 536 us (below 512 KB the blocks' fixed costs dominate;
// above it the workgroups, which drift apart by a few blocks, gather from
// more x than an L2 holds).
constexpr int64_t kColBlockEngage = 4 << 20;  // x bytes above which blocks pay
constexpr int kColBlockCols = 1 << 16;
constexpr int kColBlockMax = 1024;
constexpr int kPanelTarget = 512;  // panels: two workgroups per CU

// Panels of consecutive rows with about nnz / kPanelTarget entries each and
// at most kPanelRows rows; per panel and block the entries in CSR order, cut
// into kPanelWaves chunks at row boundaries (about equal entries each),
// every chunk padded to a whole number of quads.
void build_panels(Mat &M, int nb, int width) {
  const int64_t target = std::max<int64_t>(1, (M.nnz + kPanelTarget - 1) / kPanelTarget);
  std::vector<int> prow{0};
  int64_t cnt = 0;
  for (int r = 0; r < M.rows; ++r) {
    if (r > prow.back() && (r - prow.back() >= kPanelRows || cnt >= target)) {
      prow.push_back(r);
      cnt = 0;
    }
    cnt += M.h_ai[r + 1] - M.h_ai[r];
  }
  prow.push_back(M.rows);
  const int np = static_cast<int>(prow.size()) - 1;
  const int nc = nb * (kPanelWaves + 1);
  std::vector<std::vector<unsigned>> pi(np);
  std::vector<std::vector<int>> ps(np), pc(np);
#pragma omp parallel for schedule(dynamic, 4)
  for (int p = 0; p < np; ++p) {
    const int ra = prow[p], rb = prow[p + 1];
    std::vector<std::vector<std::pair<unsigned, int>>> lists(nb);
    for (int r = ra; r < rb; ++r)
      for (int q = M.h_ai[r]; q < M.h_ai[r + 1]; ++q) {
        const int c = M.h_aj[q], b = c / width;
        lists[b].push_back({(static_cast<unsigned>(r - ra) << kPanelColBits) |
                                static_cast<unsigned>(c - b * width),
                            q});
      }
    auto &I = pi[p];
    auto &S = ps[p];
    auto &C = pc[p];
    C.assign(nc, 0);
    for (int b = 0; b < nb; ++b) {
      const auto &L = lists[b];
      const int64_t m = static_cast<int64_t>(L.size());
      int64_t pos = 0;
      for (int w = 0; w < kPanelWaves; ++w) {
        C[b * (kPanelWaves + 1) + w] = static_cast<int>(I.size());
        int64_t e = m;
        if (w + 1 < kPanelWaves) {
          e = std::max<int64_t>(pos, (w + 1) * m / kPanelWaves);
          while (e > pos && e < m && (L[e].first >> kPanelColBits) == (L[e - 1].first >> kPanelColBits))
            ++e;
        }
        for (int64_t t = pos; t < e; ++t) {
          I.push_back(L[t].first);
          S.push_back(L[t].second);
        }
        while (I.size() % 4) {
          I.push_back(kPanelPad);
          S.push_back(-1);
        }
        pos = e;
      }
      C[b * (kPanelWaves + 1) + kPanelWaves] = static_cast<int>(I.size());
    }
  }
  std::vector<int64_t> base(np + 1, 0);
  for (int p = 0; p < np; ++p) base[p + 1] = base[p] + static_cast<int64_t>(pi[p].size());
  const int64_t len = base[np];
  if (len >= INT32_MAX) throw la_error(LA_ERR_ALLOC, "panel layout exceeds int32 offsets");
  std::vector<unsigned> gi(static_cast<size_t>(len));
  std::vector<int> gs(static_cast<size_t>(len)), gc(static_cast<size_t>(np) * nc);
#pragma omp parallel for schedule(dynamic, 16)
  for (int p = 0; p < np; ++p) {
    std::copy(pi[p].begin(), pi[p].end(), gi.begin() + base[p]);
    std::copy(ps[p].begin(), ps[p].end(), gs.begin() + base[p]);
    for (int k = 0; k < nc; ++k)
      gc[static_cast<size_t>(p) * nc + k] = static_cast<int>(base[p] + pc[p][k]);
  }
  M.pb_np = np;
  M.pb_nb = nb;
  M.pb_width = width;
  M.pb_len = len;
  M.pb_prow.alloc(sizeof(int) * prow.size());
  M.pb_prow.upload(prow.data(), sizeof(int) * prow.size(), M.stream);
  M.pb_chunk.alloc(sizeof(int) * gc.size());
  M.pb_chunk.upload(gc.data(), sizeof(int) * gc.size(), M.stream);
  M.pb_idx.alloc(sizeof(unsigned) * std::max<int64_t>(len, 4));
  M.pb_idx.upload(gi.data(), sizeof(unsigned) * len, M.stream);
  M.pb_src.alloc(sizeof(int) * std::max<int64_t>(len, 4));
  M.pb_src.upload(gs.data(), sizeof(int) * len, M.stream);
  LA_HIP(hipStreamSynchronize(M.stream));
}

bool prepare_colblock(Mat &M) {
  if (M.cb_state) return M.cb_state > 0;
  M.cb_state = -1;
  const int64_t xbytes = static_cast<int64_t>(M.cols) * 8;
  if (xbytes <= kColBlockEngage || M.nnz == 0) return false;
  const int nb = static_cast<int>((M.cols + kColBlockCols - 1) / kColBlockCols);
  if (nb > kColBlockMax) return false;
  const int width = static_cast<int>((M.cols + nb - 1) / nb);
  // scattered: most columns farther from their row than a quarter of the
  // 4 MB reach (a banded matrix keeps its gathers near the row: one pass
  // is better)
  const int64_t reach = kColBlockEngage / 8;
  int64_t far = 0;
  for (int r = 0; r < M.rows; ++r)
    for (int q = M.h_ai[r]; q < M.h_ai[r + 1]; ++q)
      far += std::llabs(static_cast<int64_t>(M.h_aj[q]) - r) > reach / 4;
  if (far * 2 < M.nnz) return false;
  build_panels(M, nb, width);
  M.pb_ver = ~0ull;
  M.cb_state = 1;
  return true;
}

// the panel layout's values follow A's (every stored precision)
void colblock_values(Mat &M, la_dtype val, hipStream_t s) {
  if (M.pb_ver != M.vals.version) {
    const unsigned g = static_cast<unsigned>((M.pb_len + 255) / 256);
    const double *av = static_cast<const double *>(M.vals.get(LA_F64));
    if (M.pb_val.has[LA_F64]) {
      double *pv = M.pb_val.v[LA_F64].as<double>();
      hipLaunchKernelGGL(k_gather_vals, dim3(g), dim3(256), 0, s, M.pb_len,
                         M.pb_src.as<int>(), av, pv);
      LA_HIP(hipGetLastError());
      M.pb_val.set_from_device_f64(pv, M.pb_len, s);
    } else {
      DevBuf t;
      t.alloc(sizeof(double) * std::max<int64_t>(M.pb_len, 1));
      hipLaunchKernelGGL(k_gather_vals, dim3(g), dim3(256), 0, s, M.pb_len,
                         M.pb_src.as<int>(), av, t.as<double>());
      LA_HIP(hipGetLastError());
      M.pb_val.set_from_device_f64(t.as<double>(), M.pb_len, s);
      LA_HIP(hipStreamSynchronize(s));
    }
    M.pb_ver = M.vals.version;
  }
  if (!M.pb_val.has[val]) M.pb_val.add(val, s);
}
}  // namespace

la_spmv_alg spmv_resolve(Mat &M, la_spmv_alg alg) {
  if (M.nnz == 0 || M.cols == 0) return LA_SPMV_ROW;
  if (alg == LA_SPMV_AUTO) {
    if (prepare_sell(M)) return LA_SPMV_SELL;
    return prepare_colblock(M) ? LA_SPMV_COLBLOCK : LA_SPMV_ALBUS;
  }
  if (alg == LA_SPMV_SELL) return prepare_sell(M) ? LA_SPMV_SELL : LA_SPMV_ALBUS;
  if (alg == LA_SPMV_COLBLOCK) return prepare_colblock(M) ? LA_SPMV_COLBLOCK : LA_SPMV_ALBUS;
  return alg == LA_SPMV_SEGSUM ? LA_SPMV_ALBUS : alg;
}

void spmv_prepare_segments(Mat &A) {
  const int64_t nseg = std::max<int64_t>(1, (A.nnz + kSegNnz - 1) / kSegNnz);
  A.nseg = nseg;
  A.seg_first_row.alloc(sizeof(int) * (nseg + 1));
  A.seg_head.alloc(sizeof(double) * nseg);
  A.seg_tail.alloc(sizeof(double) * nseg);
  A.seg_tail_row.alloc(sizeof(int) * nseg);
  A.seg_tail_end.alloc(sizeof(int) * nseg);
  hipLaunchKernelGGL(k_seg_table, dim3((nseg + 1 + 255) / 256), dim3(256), 0, A.stream, A.rows,
                     nseg, A.ai.as<int>(), A.seg_first_row.as<int>());
  LA_HIP(hipGetLastError());
  LA_HIP(hipStreamSynchronize(A.stream));
  A.seg_ready = true;
}

void spmv(Mat &A, la_spmv_alg alg, la_dtype val, la_dtype vec, la_dtype acc, double alpha,
          const void *x, double beta, const void *z, void *y, hipStream_t s) {
  LA_CHECK_ARG(val >= LA_F64 && val <= LA_BF16, "bad value dtype");
  LA_CHECK_ARG(vec == LA_F64 || vec == LA_F32, "vectors must be FP64 or FP32");
  LA_CHECK_ARG(acc == LA_F64 || acc == LA_F32, "accumulation must be FP64 or FP32");
  LA_CHECK_ARG(alg >= LA_SPMV_ROW && alg <= LA_SPMV_COLBLOCK, "bad SpMV algorithm");
  LA_CHECK_ARG(A.rows == 0 || (x && y), "null vector");
  if (A.rows == 0) return;
  if ((alg == LA_SPMV_AUTO || alg == LA_SPMV_COLBLOCK) && A.nnz && A.cols &&
      spmv_resolve(A, alg) == LA_SPMV_COLBLOCK) {
    // y = alpha A_0 x + beta z, then y += alpha A_b x for the other blocks
    colblock_values(A, val, s);
    const void *pv = A.pb_val.get(val);
    if (acc == LA_F64) {
      LA_CHECK_ARG(val == LA_F64 && vec == LA_F64,
                   "FP64 accumulation needs FP64 values and vectors");
      launch_panel<double, double, double>(A, pv, x, y, z, alpha, beta, s);
    } else if (vec == LA_F64) {
      if (val == LA_F64) launch_panel<double, double, float>(A, pv, x, y, z, alpha, beta, s);
      else if (val == LA_F32) launch_panel<float, double, float>(A, pv, x, y, z, alpha, beta, s);
      else launch_panel<bf16, double, float>(A, pv, x, y, z, alpha, beta, s);
    } else {
      if (val == LA_F64) launch_panel<double, float, float>(A, pv, x, y, z, alpha, beta, s);
      else if (val == LA_F32) launch_panel<float, float, float>(A, pv, x, y, z, alpha, beta, s);
      else launch_panel<bf16, float, float>(A, pv, x, y, z, alpha, beta, s);
    }
    return;
  }
  if (alg == LA_SPMV_COLBLOCK) alg = LA_SPMV_ALBUS;
  if (!A.vals.has[val]) A.vals.add(val, s);
  const void *av = A.vals.get(val);
  if (acc == LA_F64) {
    LA_CHECK_ARG(val == LA_F64 && vec == LA_F64,
                 "FP64 accumulation needs FP64 values and vectors");
    dispatch_alg<double, double, double>(A, alg, av, x, y, z, alpha, beta, s);
    return;
  }
  // FP32 accumulation
  if (vec == LA_F64) {
    if (val == LA_F64) dispatch_alg<double, double, float>(A, alg, av, x, y, z, alpha, beta, s);
    else if (val == LA_F32) dispatch_alg<float, double, float>(A, alg, av, x, y, z, alpha, beta, s);
    else dispatch_alg<bf16, double, float>(A, alg, av, x, y, z, alpha, beta, s);
  } else {
    if (val == LA_F64) dispatch_alg<double, float, float>(A, alg, av, x, y, z, alpha, beta, s);
    else if (val == LA_F32) dispatch_alg<float, float, float>(A, alg, av, x, y, z, alpha, beta, s);
    else dispatch_alg<bf16, float, float>(A, alg, av, x, y, z, alpha, beta, s);
  }
}

namespace {
// values of A^T from A's: out[t] = v[src[t]]
__global__ void k_gather_values(int64_t m, const int *__restrict__ src,
                                const double *__restrict__ v, double *out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < m) out[t] = v[src[t]];
}
}  // namespace

Mat &ensure_transposed(Mat &A, hipStream_t s) {
  if (!A.transposed) {
    // the pattern of A^T and, for each of its entries, the entry of A it
    // holds (a stable transpose of the entry indices), once per pattern;
    // the values follow on the device
    std::vector<int> ti(static_cast<size_t>(A.cols) + 1), tj(static_cast<size_t>(A.nnz));
    std::vector<double> idx(static_cast<size_t>(A.nnz)), tidx(static_cast<size_t>(A.nnz));
    for (int64_t q = 0; q < A.nnz; ++q) idx[q] = static_cast<double>(q);  // exact below 2^53
    transpose(A.rows, A.cols, 0, A.h_ai.data(), A.h_aj.data(), idx.data(), ti.data(), tj.data(),
              tidx.data());
    std::vector<int> src(tidx.begin(), tidx.end());
    std::vector<double> zeros(static_cast<size_t>(A.nnz), 0.0);
    A.transposed = make_mat(A.cols, A.rows, 0, ti.data(), tj.data(), zeros.data(), s);
    A.t_src.alloc(sizeof(int) * std::max<size_t>(src.size(), 1));
    A.t_src.upload(src.data(), sizeof(int) * src.size(), s);
    A.t_ver = ~0ull;
  }
  Mat &T = *A.transposed;
  T.stream = A.stream;
  if (A.t_ver != A.vals.version) {
    if (A.nnz) {
      double *tv = T.vals.v[LA_F64].as<double>();
      hipLaunchKernelGGL(k_gather_values, dim3(static_cast<unsigned>((A.nnz + 255) / 256)),
                         dim3(256), 0, s, A.nnz, A.t_src.as<int>(),
                         static_cast<const double *>(A.vals.get(LA_F64)), tv);
      LA_HIP(hipGetLastError());
      T.vals.set_from_device_f64(tv, A.nnz, s);
    }
    A.t_ver = A.vals.version;
  }
  for (int dt = 1; dt < 3; ++dt)
    if (A.vals.has[dt] && !T.vals.has[dt]) T.vals.add(static_cast<la_dtype>(dt), s);
  return T;
}

bool mat_symmetric_pattern(Mat &A) {
  if (A.sym_pattern < 0)
    A.sym_pattern = A.rows == A.cols && pattern_symmetric(A.rows, 0, A.h_ai.data(), A.h_aj.data())
                        ? 1 : 0;
  return A.sym_pattern == 1;
}

void spmv_transpose(Mat &A, la_dtype val, la_dtype vec, double alpha, const void *x,
                    double beta, void *y, hipStream_t s) {
  Mat &T = ensure_transposed(A, s);
  const la_dtype acc = (val == LA_F64 && vec == LA_F64) ? LA_F64 : LA_F32;
  spmv(T, LA_SPMV_AUTO, val, vec, acc, alpha, x, beta, y, y, s);
}

void csr_value_grad(Mat &A, const double *lambda, const double *x, double *g, hipStream_t s) {
  if (A.rows == 0) return;
  hipLaunchKernelGGL(k_value_grad, dim3((A.rows + 255) / 256), dim3(256), 0, s, A.rows,
                     A.ai.as<int>(), A.aj.as<int>(), lambda, x, g);
  LA_HIP(hipGetLastError());
}

}  // namespace laamd
