// Sparse triangular solves on gfx950 and the ILU(k)/IC(k) apply.
//
// Analysis (host): level sets by TopologicalSort2
// (sparse_mat_op/matrix_utils.hpp:637-675); the triangle is stored with its
// rows in level order (the reference's permuteRow in
// OptimizedTriangularSolve::analysis, triangle_solve.tpp:87-91) so each
// level is a contiguous, coalesced stream; columns keep original ids.
//
// Solve, two modes:
//  * LEVEL: one launch per level (the Barrier mode, triangle_solve.tpp:
//    127-155), captured once into a HIP graph and replayed.
//  * SYNCFREE: one launch; workgroups claim chunks of rows in level order
//    from an atomic ticket (the NoBarrier mode, triangle_solve.tpp:157-187).
//    x is pre-filled with a sentinel; a row waits until each x_j it reads is
//    no longer the sentinel (value and readiness in one word) and publishes
//    its own x_i the same way. Chunks never span levels, so rows inside one
//    workgroup never wait on each other, and every wait is on an earlier
//    ticket, which a running workgroup already holds: the solve always
//    drains. Waits are bounded; a timeout sets an error flag instead of
//    hanging.
//  * JACOBI (no counterpart in the reference): k Jacobi sweeps on the
//    triangle, y_0 = D^-1 b, y_{s+1} = D^-1 (b - N y_s), each sweep an
//    SpMV with the column-scaled strict triangle N D^-1. D^-1 N is nilpotent (its
//    index is the number of levels), so k >= levels - 1 sweeps give the
//    exact solution; fewer give the truncated Neumann series, an
//    approximate solve whose cost is k bandwidth-bound passes instead of a
//    chain of level-to-level dependencies (Anzt, Chow, Dongarra, "Iterative
//    sparse triangular solves for preconditioning", Euro-Par 2015). For an
//    IC factor both triangles use the same k, so the applied operator
//    P_k^T P_k stays symmetric positive definite and PCG stays valid.
// Row formula (triangle_solve.hpp:32-71): x_i = (b_i - sum_j T_ij x_j) / d_i,
// computed as a product with 1/d_i stored in the plan (within an ulp of the
// quotient; keeps a divide off the sync-free dependency chain).
#include <algorithm>
#include <cstdlib>

#include "dev_util.hpp"
#include "device.hpp"
#include "host_algos.hpp"

namespace laamd {
using namespace dev;

namespace {
constexpr int kB = 256;

template <typename T>
__device__ __forceinline__ auto ldt(const T *p, int64_t i) {
  return ldm(p, i);
}

template <int G, typename T, typename V, typename A, bool DIAG>
__global__ __launch_bounds__(kB) void k_trsv_level(int k0, int k1, const int *__restrict__ rowptr,
                                                   const int *__restrict__ col,
                                                   const T *__restrict__ val,
                                                   const T *__restrict__ diag,
                                                   const int *__restrict__ perm,
                                                   const V *__restrict__ b, V *x) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  const int64_t k = k0 + t / G;
  const int sub = threadIdx.x & (G - 1);
  if (k >= k1) return;
  const int rs = rowptr[k], re = rowptr[k + 1];
  A acc = 0;
  for (int q = rs + sub; q < re; q += G) acc += A(ldt(val, q)) * A(x[ldi(col, q)]);
  acc = group_sum<G>(acc);
  if (sub == 0) {
    const int r = perm[k];
    A v = A(b[r]) - acc;
    if constexpr (DIAG) v = v * A(ldt(diag, k));
    x[r] = V(v);
  }
}

// x <- sentinel everywhere; the control words ctl[0..15] (ctl[0]: ticket,
// ctl[1]: owning XCD + 1) reset for the next sync-free solve.
template <typename V>
__global__ void k_fill_sentinel(int64_t n, V *x, unsigned *ctl) {
  using U = typename Sentinel<V>::U;
  U *xb = reinterpret_cast<U *>(x);
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB)
    xb[t] = Sentinel<V>::bits;
  if (blockIdx.x == 0 && threadIdx.x < 16) ctl[threadIdx.x] = 0;
}

// Sync-free solve. Readiness is polled with agent-scope relaxed loads (they
// bypass the CU's L1).
//  XLOCAL = true: confined to one XCD. The first workgroup to start claims
//    its XCD (device-coherent CAS on ctl[1]); workgroups that land on another
//    XCD leave at once. Results are published with plain stores into that
//    XCD's L2, where the agent-scope polls of the same XCD see them: about
//    0.3 us from store to poll on MI355X, against 0.55 us when producer and
//    consumer sit on different XCDs.
//  XLOCAL = false: every XCD takes part; results are published with
//    agent-scope stores so polls from any XCD see them.
// A workgroup claims one chunk ticket from ctl[0] at a time (one device
// atomic per NT/G rows; the tickets of one counter are handed out about
// 12 ns apart, so a per-wavefront ticket would bound wide levels). Chunks
// are in level order. A wait longer than kWaitTicks of the 100 MHz
// real-time clock sets *err and gives up; every other waiter checks *err
// every 256 polls and gives up too, so the kernel always drains quickly.
constexpr uint64_t kWaitTicks = 20000000;  // 0.2 s

// One row of a sync-free solve, G lanes: wait for each x_j, accumulate,
// publish x_r. `sub` is the lane's index inside its group.
template <int G, typename T, typename V, typename A, bool DIAG, bool XLOCAL>
__device__ __forceinline__ void sf_row(int k, int sub, const int *__restrict__ rowptr,
                                       const int *__restrict__ col, const T *__restrict__ val,
                                       const T *__restrict__ diag, const int *__restrict__ perm,
                                       const V *__restrict__ b, V *x, int *err) {
  using U = typename Sentinel<V>::U;
  // everything that does not depend on x first
  const int rs = rowptr[k], re = rowptr[k + 1];
  const int r = perm[k];
  const A bb = A(b[r]);
  A dd = A(1);
  if constexpr (DIAG) dd = A(ldt(diag, k));
  A acc = 0;
  bool gave_up = false;
  // one entry per lane at a time: a lane that loads only entries of its own
  // row issues no extra request (the polls below go to the memory side, and
  // their count is what the solve pays for)
  for (int q = rs + sub; q < re; q += G) {
    const int j = ldi(col, q);
    const A a = A(ldt(val, q));
    V xv = load_ready(x + j);
    if (is_sentinel(xv) && !gave_up) {
      const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
      unsigned polls = 0;
      while (is_sentinel(xv)) {
        __builtin_amdgcn_s_sleep(1);
        xv = load_ready(x + j);
        if ((++polls & 255u) == 0) {
          if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
              __builtin_amdgcn_s_memrealtime() - t0 > kWaitTicks) {
            gave_up = true;
            break;
          }
        }
      }
    }
    acc += a * A(is_sentinel(xv) ? V(0) : xv);
  }
  if (gave_up) __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  acc = group_sum<G>(acc);
  if (sub == 0) {
    A v = bb - acc;
    if constexpr (DIAG) v = v * dd;
    U o = to_bits<V>(V(v));
    if (o == Sentinel<V>::bits) o = Sentinel<V>::qnan;
    if constexpr (XLOCAL)
      __hip_atomic_store(reinterpret_cast<U *>(x + r), o, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
    else
      __hip_atomic_store(reinterpret_cast<U *>(x + r), o, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <int G, int NT, typename T, typename V, typename A, bool DIAG, bool XLOCAL>
__global__ __launch_bounds__(NT) void k_trsv_syncfree(
    int nchunks, const int *__restrict__ chunk_start, const int *__restrict__ rowptr,
    const int *__restrict__ col, const T *__restrict__ val, const T *__restrict__ diag,
    const int *__restrict__ perm, const V *__restrict__ b, V *x, unsigned *ctl, int *err) {
  if constexpr (XLOCAL) {
    __shared__ int s_go;
    if (threadIdx.x == 0) {
      const unsigned me = xcc_id() + 1;
      const unsigned prev = atomicCAS(&ctl[1], 0u, me);
      s_go = (prev == 0u || prev == me) ? 1 : 0;
    }
    __syncthreads();
    if (!s_go) return;
  }
  // row of this lane group inside a chunk of NT/G rows
  const int gi = static_cast<int>(threadIdx.x) / G, sub = threadIdx.x & (G - 1);
  // two ticket slots: while a chunk is solved, thread 0 already claims the
  // next ticket into the other slot, so the claim's round trip overlaps the
  // work; one barrier per chunk. Holding one ticket ahead keeps the drain
  // argument: the lowest unfinished chunk is either some workgroup's
  // current chunk or the next one of a workgroup whose current chunk is
  // finished.
  __shared__ int s_c[2];
  if (threadIdx.x == 0) s_c[0] = static_cast<int>(atomicAdd(&ctl[0], 1u));
  __syncthreads();
  for (int it = 0;; ++it) {
    const int c = s_c[it & 1];
    if (c >= nchunks) break;
    if (threadIdx.x == 0) s_c[(it + 1) & 1] = static_cast<int>(atomicAdd(&ctl[0], 1u));
    const int k = chunk_start[c] + gi;
    if (k < chunk_start[c + 1])
      sf_row<G, T, V, A, DIAG, XLOCAL>(k, sub, rowptr, col, val, diag, perm, b, x, err);
    __syncthreads();
  }
}

// ------------------------------------------------------- tiled schedule
// One workgroup per tile (see tiling.cpp): wavefront 0 solves the tile's
// groups (at most 64 rows of one level each) in level order, one lane per
// row, reading every x_j from LDS: its own results of the last kTileWindow
// groups or the values of other tiles (the halo), through one 16-bit LDS
// index per entry. The rows' data is stored group by group, 64 slots per
// group, and the solving wavefront loads it straight from HBM into
// registers: each group's indices, values and 1/d, and b (gathered by row
// id) kTileDepth groups ahead of the group it solves, the row ids twice as
// far ahead. A group's dependency chain is then the x reads from LDS, the
// products, and the LDS write and HBM store of the results (padding lanes
// store nothing). Wavefronts 1..H fetch the other tiles' values:
// wavefront w takes the tile's groups w-1, w-1+H, ..., polls the
// values each group reads first (agent-scope loads of x, as in the
// level-chunk kernel) into the halo, loading the next group's list while
// it polls, and counts its finished groups in `s_prog[w-1]`; the solver
// checks that count before a group's x reads. Halo slots are never reused
// within a tile, so the fetching wavefronts run ahead freely. A wavefront's
// LDS accesses are carried out in the order it issues them, so the solving
// wavefront needs no wait after its own writes (a compiler-only fence keeps
// the order in the code); the fetching wavefronts wait for their LDS writes
// before counting a group. Every wait is bounded like the level-chunk
// kernel's; a timeout sets *err.
//
// H, the number of fetching wavefronts, trades halo throughput against the
// solving wavefront's share of its SIMD and the memory traffic of the
// polls: with 7 or 8 wavefronts most SIMDs hold two, and the solver may
// share issue with a fetcher; with 4, the solver has its SIMD to itself.
// Rows on one lane (2-D and 7-point tiles, up to 64 rows and many halo
// values per group) take 6; rows on kTileRowLanes lanes (box tiles, 16 rows
// per group) take 3.
constexpr int kTileLanes = 64;
constexpr int kTileHalo = 6;      // fetching wavefronts for rows on one lane
constexpr int kTileHaloWide = 3;  // for rows on kTileRowLanes lanes
constexpr int kTileDepth = 8;   // groups whose rows sit in registers
constexpr int kTileXl = kTileWindow * kTileLanes;  // result slots; kTileXl is the zero slot

__device__ __forceinline__ void lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup", "local");
}
__device__ __forceinline__ void code_order() { __atomic_signal_fence(__ATOMIC_SEQ_CST); }
__device__ __forceinline__ int lds_load(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_store(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// N bytes from p (16-byte aligned for N >= 16, 8-byte for N == 8) into out
typedef unsigned int tile_u2 __attribute__((ext_vector_type(2)));
typedef unsigned int tile_u4 __attribute__((ext_vector_type(4)));
template <int N>
__device__ __forceinline__ void ld_bytes(const void *p, void *out) {
  static_assert(N == 8 || N % 16 == 0, "load size");
  if constexpr (N == 8) {
    *reinterpret_cast<tile_u2 *>(out) =
        __builtin_nontemporal_load(reinterpret_cast<const tile_u2 *>(p));
  } else {
#pragma unroll
    for (int q = 0; q < N / 16; ++q)
      reinterpret_cast<tile_u4 *>(out)[q] =
          __builtin_nontemporal_load(reinterpret_cast<const tile_u4 *>(p) + q);
  }
}

template <typename T> struct TileDt;
template <> struct TileDt<double> { static constexpr la_dtype v = LA_F64; };
template <> struct TileDt<float> { static constexpr la_dtype v = LA_F32; };
template <> struct TileDt<bf16> { static constexpr la_dtype v = LA_BF16; };

template <typename T> __device__ __forceinline__ float tile_val(T v) { return float(v); }
template <> __device__ __forceinline__ float tile_val<bf16>(bf16 v) {
  return __uint_as_float(static_cast<uint32_t>(v.bits) << 16);
}
__device__ __forceinline__ double tile_val(double v) { return v; }

// one pipeline stage: a group's slot for this lane
template <int E, typename T>
struct TileStage {
  alignas(16) uint16_t code[E];
  alignas(16) T val[E];
  T dinv;
};

template <int E, int RL, int H, typename T, typename V, typename A, bool DIAG>
__global__ __launch_bounds__(kTileLanes * (1 + H)) void k_trsv_tiled(
    int nparts, const int *__restrict__ gptr, const int *__restrict__ prow,
    const uint16_t *__restrict__ pcode, const T *__restrict__ pval, const T *__restrict__ pdinv,
    const int *__restrict__ rfptr, const int *__restrict__ rfcol, const int *__restrict__ rfslot,
    const V *__restrict__ b, V *x, unsigned *ctl, int *err) {
  // pipeline depth: slots of 8 entries hold twice the registers
  constexpr int D = E <= 4 ? kTileDepth : kTileDepth * 3 / 4;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_prog[H];
  __shared__ int s_part, s_late;
  V *xs = reinterpret_cast<V *>(smem);  // window, zero slot, halo
  const int wave = static_cast<int>(threadIdx.x) / kTileLanes;
  const int lane = static_cast<int>(threadIdx.x) % kTileLanes;
  for (;;) {
    if (threadIdx.x == 0) {
      s_part = static_cast<int>(atomicAdd(&ctl[0], 1u));
      s_late = 0;
      xs[kTileXl] = V(0);
    }
    if (threadIdx.x < H) s_prog[threadIdx.x] = 0;
    __syncthreads();
    const int p = s_part;
    if (p >= nparts) break;
    const int g0 = gptr[p], ng = gptr[p + 1] - g0;
    if (wave == 0 && ng > 0) {
      __builtin_amdgcn_s_setprio(3);  // issue priority over the halo wavefronts
      // group j's slot of this lane (j clamped: loads past the last group
      // repeat it, so every iteration issues the same loads)
      auto slot_of = [&](int j) {
        return static_cast<int64_t>(g0 + min(j, ng - 1)) * kTileLanes + lane;
      };
      // row ids 2D groups ahead, the rows' data and b (gathered by row id)
      // D groups ahead
      int rid[2 * D];
      TileStage<E, T> st[D];
      V bv[D];
      auto load_rows = [&](int k, int j, int rk) {  // rk: the ring slot of j's row id
        const int64_t sl = slot_of(j);
        ld_bytes<E * 2>(pcode + sl * E, st[k].code);
        ld_bytes<E * static_cast<int>(sizeof(T))>(pval + sl * E, st[k].val);
        if constexpr (DIAG) st[k].dinv = pdinv[sl];
        bv[k] = b[max(rid[rk], 0)];
      };
      // prime in the order the loop keeps, so that the loads in flight
      // when the loop starts are as old as in any later pass (the
      // compiler's count of loads allowed in flight then stays high;
      // sched_barrier keeps the scheduler from reordering them)
#pragma unroll
      for (int k = 0; k < 2 * D; ++k) rid[k] = __builtin_nontemporal_load(prow + slot_of(k));
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < D; ++k) {
        load_rows(k, k, k);
        __builtin_amdgcn_sched_barrier(0);
      }
      int rdy = lds_load(&s_prog[0]);
      for (int i0 = 0; i0 < ng; i0 += 2 * D) {
#pragma unroll
        for (int k2 = 0; k2 < 2 * D; ++k2) {
          // groups past the last one (at most 2D - 1) repeat its work
          // without publishing: a loop without early exits keeps the
          // compiler's count of loads in flight across iterations
          const int i = i0 + k2, k = k2 % D;
          const bool real = i < ng;
          // the halo values of group i are in LDS once its fetcher has
          // counted it
          if (real && rdy <= i / H) {
            // no global access on this path: the loads in flight are
            // counted in issue order, and a global access here would make
            // the compiler wait for all of them
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            for (unsigned polls = 1; lds_load(&s_prog[i % H]) <= i / H; ++polls) {
              __builtin_amdgcn_s_sleep(0);
              if ((polls & 1023u) == 0 && __builtin_amdgcn_s_memrealtime() - t0 > kWaitTicks) {
                s_late = 1;
                break;
              }
            }
            lds_fence();
          }
          code_order();
          V xv[E];
#pragma unroll
          for (int e = 0; e < E; ++e) xv[e] = xs[st[k].code[e]];
          A acc = A(tile_val(st[k].val[0])) * A(xv[0]);
#pragma unroll
          for (int e = 1; e < E; ++e) acc += A(tile_val(st[k].val[e])) * A(xv[e]);
          // a row on RL lanes: every lane gets the same sum (the first
          // lane's result is the one read and published)
          acc = group_sum<RL>(acc);
          A v = A(bv[k]) - acc;
          if constexpr (DIAG) v = v * A(tile_val(st[k].dinv));
          const V out = V(v);
          const int r = rid[k2];
          xs[(i % kTileWindow) * kTileLanes + lane] = out;
          if (real && r >= 0) publish(x + r, out);
          code_order();
          rdy = lds_load(&s_prog[(i + 1) % H]);
          // refill: the row id of group i + 2D, the rows of group i + D
          rid[k2] = __builtin_nontemporal_load(prow + slot_of(i + 2 * D));
          load_rows(k, i + D, (k2 + D) % (2 * D));
        }
      }
    } else if (wave > 0) {
      // this wavefront's groups j = hw, hw + H, ... (hw = wave - 1): the
      // next group's list (its halo sources and slots) is loaded while the
      // current group's values are polled
      const int hw = wave - 1;
      int j = hw, f0 = 0, f1 = 0, col = 0, slot = 0;
      if (j < ng) {
        f0 = rfptr[g0 + j];
        f1 = rfptr[g0 + j + 1];
        if (f0 + lane < f1) {
          col = rfcol[f0 + lane];
          slot = rfslot[f0 + lane];
        }
      }
      for (; j < ng; j += H) {
        V xv = V(0);
        if (f0 + lane < f1) xv = load_ready(x + col);
        const int jn = j + H;
        int nf0 = 0, nf1 = 0, ncol = 0, nslot = 0;
        if (jn < ng) {
          nf0 = rfptr[g0 + jn];
          nf1 = rfptr[g0 + jn + 1];
          if (nf0 + lane < nf1) {
            ncol = rfcol[nf0 + lane];
            nslot = rfslot[nf0 + lane];
          }
        }
        for (int t = f0 + lane; t < f1; t += kTileLanes) {
          const int c = t == f0 + lane ? col : rfcol[t];
          if (t != f0 + lane) xv = load_ready(x + c);
          if (is_sentinel(xv)) {
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            for (unsigned polls = 1; is_sentinel(xv); ++polls) {
              __builtin_amdgcn_s_sleep(1);
              xv = load_ready(x + c);
              if ((polls & 255u) == 0 &&
                  (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                   __builtin_amdgcn_s_memrealtime() - t0 > kWaitTicks)) {
                __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                xv = V(0);
                break;
              }
            }
          }
          xs[t == f0 + lane ? slot : rfslot[t]] = xv;
        }
        lds_fence();
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) lds_store(&s_prog[hw], j / H + 1);
        f0 = nf0;
        f1 = nf1;
        col = ncol;
        slot = nslot;
      }
    }
    __syncthreads();
    // a solver wait that ran out of time is reported here
    if (threadIdx.x == 0 && s_late)
      __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

inline int pick_G(double avg) {
  int G = 1;
  while (G < 64 && G < avg) G <<= 1;
  return G;
}

template <typename T, typename V, typename A, bool DIAG>
void launch_level_range(const TrsvPlan &P, int l0, int l1, const T *val, const T *diag,
                        const V *b, V *x, hipStream_t s) {
  for (int l = l0; l < l1; ++l) {
    const int k0 = P.level_ptr[l], k1 = P.level_ptr[l + 1];
    if (k1 <= k0) continue;
    const int64_t threads = static_cast<int64_t>(k1 - k0) * P.G;
    const unsigned grid = static_cast<unsigned>((threads + kB - 1) / kB);
#define LA_LVL(GG)                                                                           \
  hipLaunchKernelGGL((k_trsv_level<GG, T, V, A, DIAG>), dim3(grid), dim3(kB), 0, s, k0, k1, \
                     P.rowptr.as<int>(), P.col.as<int>(), val, diag, P.perm.as<int>(), b, x)
    switch (P.G) {
      case 1: LA_LVL(1); break;
      case 2: LA_LVL(2); break;
      case 4: LA_LVL(4); break;
      case 8: LA_LVL(8); break;
      case 16: LA_LVL(16); break;
      case 32: LA_LVL(32); break;
      default: LA_LVL(64); break;
    }
#undef LA_LVL
  }
  LA_HIP(hipGetLastError());
}

template <int G, int NT, bool XL, typename T, typename V, typename A, bool DIAG>
void sf_launch(TrsvPlan &P, const T *val, const T *diag, const V *b, V *x, hipStream_t s) {
  int *rp = P.rowptr.as<int>(), *cl = P.col.as<int>(), *pm = P.perm.as<int>();
  unsigned *ctl = P.counter.as<unsigned>();
  int *err = P.err_flag.as<int>();
  // XLOCAL launches 8x the workgroups it needs (the others exit at once);
  // otherwise the whole chip, `sf_wgs` workgroups in total
  const unsigned need = std::max(1u, static_cast<unsigned>(P.nchunks));
  const unsigned grid = XL ? 8u * std::min(need, P.sf_wgs) : std::min(need, P.sf_wgs);
  hipLaunchKernelGGL((k_trsv_syncfree<G, NT, T, V, A, DIAG, XL>), dim3(grid), dim3(NT), 0, s,
                     P.nchunks, P.chunk_start.as<int>(), rp, cl, val, diag, pm, b, x, ctl, err);
}

// x <- sentinel, control words reset (the tiled schedule's stand-in for
// k_fill_sentinel)
template <typename V>
__global__ void k_tile_prep(int64_t n, V *x, unsigned *ctl) {
  using U = typename Sentinel<V>::U;
  U *xb = reinterpret_cast<U *>(x);
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB)
    xb[t] = Sentinel<V>::bits;
  if (blockIdx.x == 0 && threadIdx.x < 16) ctl[threadIdx.x] = 0;
}

inline size_t tile_lds_bytes(int h_max, size_t vsize) {
  return (static_cast<size_t>(kTileXl) + 1 + h_max) * vsize;
}

template <int E, int RL, int H, typename T, typename V, typename A, bool DIAG>
void tiled_launch(TrsvPlan &P, const V *b, V *x, hipStream_t s) {
  auto kern = k_trsv_tiled<E, RL, H, T, V, A, DIAG>;
  const size_t lds = tile_lds_bytes(P.t_hmax, sizeof(V));
  static size_t lds_set = 0;  // per instantiation: the largest size allowed so far
  if (lds > lds_set) {
    LA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(lds)));
    lds_set = lds;
  }
  const T *tv = static_cast<const T *>(P.t_val.get(TileDt<T>::v));
  const T *td = DIAG ? static_cast<const T *>(P.t_dinv.get(TileDt<T>::v)) : nullptr;
  const unsigned fg = static_cast<unsigned>(std::min<int64_t>(4096, (P.n + kB - 1) / kB));
  hipLaunchKernelGGL(k_tile_prep<V>, dim3(std::max(fg, 1u)), dim3(kB), 0, s,
                     static_cast<int64_t>(P.n), x, P.counter.as<unsigned>());
  hipLaunchKernelGGL(kern, dim3(P.t_wgs), dim3(kTileLanes * (1 + H)), lds, s,
                     P.t_nparts, P.t_gptr.as<int>(), P.t_prow.as<int>(),
                     P.t_pcode.as<uint16_t>(), tv, td, P.t_rfptr.as<int>(), P.t_rfcol.as<int>(),
                     P.t_rfslot.as<int>(), b, x, P.counter.as<unsigned>(), P.err_flag.as<int>());
}

template <typename T, typename V, typename A, bool DIAG>
void launch_syncfree(TrsvPlan &P, const T *val, const T *diag, const V *b, V *x, hipStream_t s) {
  const int64_t n = P.n;
  if (P.tiled) {
    constexpr int E8 = kTileMaxEntries, RW = kTileRowLanes;
    if (P.t_maxe == 4) tiled_launch<4, 1, kTileHalo, T, V, A, DIAG>(P, b, x, s);
    else if (P.t_rlanes == 1) tiled_launch<E8, 1, kTileHalo, T, V, A, DIAG>(P, b, x, s);
    else tiled_launch<E8, RW, kTileHaloWide, T, V, A, DIAG>(P, b, x, s);
    LA_HIP(hipGetLastError());
    return;
  }
  unsigned fg = static_cast<unsigned>(std::min<int64_t>(4096, (n + kB - 1) / kB));
  if (fg == 0) fg = 1;
  hipLaunchKernelGGL(k_fill_sentinel<V>, dim3(fg), dim3(kB), 0, s, n, x, P.counter.as<unsigned>());
  LA_HIP(hipGetLastError());
#define LA_SF2(GG, NT)                                      \
  if (P.sf_xlocal) {                                        \
    sf_launch<GG, NT, true, T, V, A, DIAG>(P, val, diag, b, x, s);  \
  } else {                                                  \
    sf_launch<GG, NT, false, T, V, A, DIAG>(P, val, diag, b, x, s); \
  }
#define LA_SF(GG)              \
  if (P.sf_nt == 1024) {       \
    LA_SF2(GG, 1024);          \
  } else if (P.sf_nt == 512) { \
    LA_SF2(GG, 512);           \
  } else {                     \
    LA_SF2(GG, 256);           \
  }
  switch (P.G) {
    case 1: LA_SF(1); break;
    case 2: LA_SF(2); break;
    case 4: LA_SF(4); break;
    case 8: LA_SF(8); break;
    case 16: LA_SF(16); break;
    case 32: LA_SF(32); break;
    default: LA_SF(64); break;
  }
#undef LA_SF2
#undef LA_SF
  LA_HIP(hipGetLastError());
}

template <typename T, typename V, typename A>
void solve_typed(TrsvPlan &P, la_trsv_mode mode, const void *valp, const void *diagp,
                 const void *b, void *x, hipStream_t s) {
  const T *val = static_cast<const T *>(valp);
  const T *diag = static_cast<const T *>(diagp);
  const V *bb = static_cast<const V *>(b);
  V *xx = static_cast<V *>(x);
  if (mode == LA_TRSV_SYNCFREE) {
    if (P.unit) launch_syncfree<T, V, A, false>(P, val, diag, bb, xx, s);
    else launch_syncfree<T, V, A, true>(P, val, diag, bb, xx, s);
  } else {
    if (P.unit) launch_level_range<T, V, A, false>(P, 0, P.nlevels, val, diag, bb, xx, s);
    else launch_level_range<T, V, A, true>(P, 0, P.nlevels, val, diag, bb, xx, s);
  }
}

void solve_dispatch(TrsvPlan &P, la_trsv_mode mode, la_dtype val, la_dtype vec, const void *b,
                    void *x, hipStream_t s) {
  const void *v = P.nnz ? P.val.get(val) : nullptr;
  const void *d = P.unit ? nullptr : P.diag.get(val);
  if (val == LA_F64 && vec == LA_F64) return solve_typed<double, double, double>(P, mode, v, d, b, x, s);
  if (vec == LA_F64) {
    if (val == LA_F64) return solve_typed<double, double, float>(P, mode, v, d, b, x, s);
    if (val == LA_F32) return solve_typed<float, double, float>(P, mode, v, d, b, x, s);
    return solve_typed<bf16, double, float>(P, mode, v, d, b, x, s);
  }
  if (val == LA_F64) return solve_typed<double, float, float>(P, mode, v, d, b, x, s);
  if (val == LA_F32) return solve_typed<float, float, float>(P, mode, v, d, b, x, s);
  return solve_typed<bf16, float, float>(P, mode, v, d, b, x, s);
}

// ------------------------------------------------------------ JACOBI mode
// Sweeps in the column-scaled form: with w = D y, y_{s+1} = D^-1 (b - N y_s)
// becomes w_{s+1} = b - (N D^-1) w_s from w_0 = b, and x = D^-1 w_k; the
// last sweep applies D^-1 in its epilogue, so a solve is k launches (no
// separate scaling pass).

// out_i = dinv_i * b_i (k = 0, or the last sweep without a sliced-ELL form)
template <typename T, typename V, typename A>
__global__ void k_jac_scale(int n, const T *__restrict__ dinv, const V *__restrict__ b, V *out) {
  const int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = static_cast<V>(static_cast<A>(ldt(dinv, i)) * static_cast<A>(b[i]));
}

// D^-1 from a device factor array f
__global__ void k_jac_dinv(int n, const int64_t *__restrict__ dsrc, const double *__restrict__ f,
                           double *dinv) {
  const int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dinv[i] = 1.0 / f[dsrc[i]];
}

// N D^-1 from f, row by row (dinv null: unit diagonal)
__global__ void k_jac_vals(int n, const int *__restrict__ ai, const int *__restrict__ aj,
                           const int64_t *__restrict__ src, const double *__restrict__ f,
                           const double *__restrict__ dinv, double *nv) {
  const int i = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int q = ai[i]; q < ai[i + 1]; ++q) nv[q] = dinv ? f[src[q]] * dinv[aj[q]] : f[src[q]];
}

void jacobi_refresh(TrsvPlan &T, const double *f, hipStream_t s) {
  Mat &J = *T.jac;
  if (T.h_src.empty() || T.n == 0) return;
  const unsigned grid = static_cast<unsigned>((T.n + 255) / 256);
  double *dv = T.unit ? nullptr : T.jdinv.v[LA_F64].as<double>();
  if (!T.unit && !T.h_dsrc.empty()) {
    hipLaunchKernelGGL(k_jac_dinv, dim3(grid), dim3(256), 0, s, T.n, T.jdsrc.as<int64_t>(), f,
                       dv);
    LA_HIP(hipGetLastError());
    T.jdinv.set_from_device_f64(dv, T.n, s);
  }
  if (J.nnz) {
    double *nv = J.vals.v[LA_F64].as<double>();
    hipLaunchKernelGGL(k_jac_vals, dim3(grid), dim3(256), 0, s, T.n, J.ai.as<int>(),
                       J.aj.as<int>(), T.jsrc.as<int64_t>(), f, dv, nv);
    LA_HIP(hipGetLastError());
    J.vals.set_from_device_f64(nv, J.nnz, s);
  }
}

void build_jacobi(TrsvPlan &T, hipStream_t s) {
  const int n = T.n;
  std::vector<double> sv(static_cast<size_t>(T.nnz)), dv(T.unit ? 0 : n);
  for (int i = 0; i < n && !T.unit; ++i) dv[i] = 1.0 / T.h_diag[i];
  for (int i = 0; i < n; ++i)
    for (int q = T.h_ai[i]; q < T.h_ai[i + 1]; ++q)
      sv[q] = T.unit ? T.h_av[q] : T.h_av[q] * dv[T.h_aj[q]];
  T.jac = make_mat(n, n, 0, T.h_ai.data(), T.h_aj.data(), sv.data(), s);
  if (!T.unit) T.jdinv.set_from_host(dv.data(), n, s);
  T.jtmp[0].alloc(sizeof(double) * static_cast<size_t>(n));
  T.jtmp[1].alloc(sizeof(double) * static_cast<size_t>(n));
  if (!T.h_src.empty()) {
    T.jsrc.alloc(sizeof(int64_t) * std::max<size_t>(T.h_src.size(), 1));
    T.jsrc.upload(T.h_src.data(), sizeof(int64_t) * T.h_src.size(), s);
    if (!T.h_dsrc.empty()) {
      T.jdsrc.alloc(sizeof(int64_t) * T.h_dsrc.size());
      T.jdsrc.upload(T.h_dsrc.data(), sizeof(int64_t) * T.h_dsrc.size(), s);
    }
  }
  if (T.fsrc) jacobi_refresh(T, T.fsrc, s);
  LA_HIP(hipStreamSynchronize(s));
}

void jac_scale(int n, la_dtype val, la_dtype vec, const void *d, const void *b, void *o,
               hipStream_t s) {
  const dim3 grid(static_cast<unsigned>((n + 255) / 256)), blk(256);
#define LA_JS(T, V, A)                                                                          \
  hipLaunchKernelGGL((k_jac_scale<T, V, A>), grid, blk, 0, s, n, static_cast<const T *>(d),     \
                     static_cast<const V *>(b), static_cast<V *>(o))
  if (vec == LA_F64) {
    if (val == LA_F64) LA_JS(double, double, double);
    else if (val == LA_F32) LA_JS(float, double, float);
    else LA_JS(bf16, double, float);
  } else {
    if (val == LA_F64) LA_JS(double, float, float);
    else if (val == LA_F32) LA_JS(float, float, float);
    else LA_JS(bf16, float, float);
  }
#undef LA_JS
  LA_HIP(hipGetLastError());
}

void jacobi_solve(TrsvPlan &T, la_dtype val, la_dtype vec, const void *b, void *x,
                  hipStream_t s) {
  if (!T.jac) build_jacobi(T, s);
  Mat &J = *T.jac;
  const int n = T.n;
  const size_t esz = vec == LA_F64 ? 8 : 4;
  const la_dtype acc = val == LA_F64 && vec == LA_F64 ? LA_F64 : LA_F32;
  const int k = T.sweeps;
  const void *d = nullptr;
  if (!T.unit) {
    if (!T.jdinv.has[val]) T.jdinv.add(val, s);
    d = T.jdinv.get(val);
  }
  if (k == 0) {
    if (d) jac_scale(n, val, vec, d, b, x, s);
    else LA_HIP(hipMemcpyAsync(x, b, esz * n, hipMemcpyDeviceToDevice, s));
    return;
  }
  void *t0 = T.jtmp[0].p, *t1 = T.jtmp[1].p;
  // sweeps 1 .. k-1 alternate between the two spare vectors (a unit
  // triangle's sweep q writes x when k - q is even instead)
  const void *prev = b;
  for (int q = 1; q < k; ++q) {
    void *dst = d ? (q % 2 ? t0 : t1) : ((k - q) % 2 == 0 ? x : (prev == t0 ? t1 : t0));
    spmv(J, LA_SPMV_AUTO, val, vec, acc, -1.0, prev, 1.0, b, dst, s);
    prev = dst;
  }
  if (!d) {
    spmv(J, LA_SPMV_AUTO, val, vec, acc, -1.0, prev, 1.0, b, x, s);
    return;
  }
  // last sweep: x = D^-1 (b - (N D^-1) w_{k-1})
  if (spmv_sell_rowscaled(J, val, vec, d, prev, b, x, s)) return;
  void *tmp = prev == t0 ? t1 : t0;
  spmv(J, LA_SPMV_AUTO, val, vec, acc, -1.0, prev, 1.0, b, tmp, s);
  jac_scale(n, val, vec, d, tmp, x, s);
}

// slot order <- level order: the plan's values (entry source index, -1 for
// padding) and reciprocal diagonal
__global__ void k_tile_gather(int64_t m, const int *__restrict__ src, const double *__restrict__ f,
                              double *out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < m) out[t] = src[t] >= 0 ? f[src[t]] : 0.0;
}

void tiled_refresh(TrsvPlan &T, hipStream_t s) {
  if (!T.tiled) return;
  const int64_t slots = static_cast<int64_t>(T.t_ngroups) * kTileLanes;
  const int64_t m = slots * T.t_maxe;
  if (T.nnz) {
    double *tv = T.t_val.v[LA_F64].as<double>();
    hipLaunchKernelGGL(k_tile_gather, dim3(static_cast<unsigned>((m + 255) / 256)), dim3(256), 0,
                       s, m, T.t_lsrc.as<int>(), T.val.v[LA_F64].as<double>(), tv);
    LA_HIP(hipGetLastError());
    T.t_val.set_from_device_f64(tv, m, s);
  }
  if (!T.unit) {
    double *td = T.t_dinv.v[LA_F64].as<double>();
    hipLaunchKernelGGL(k_tile_gather, dim3(static_cast<unsigned>((slots + 255) / 256)),
                       dim3(256), 0, s, slots, T.t_dpos.as<int>(),
                       T.diag.v[LA_F64].as<double>(), td);
    LA_HIP(hipGetLastError());
    T.t_dinv.set_from_device_f64(td, slots, s);
  }
}

// The tiled schedule when the triangle's pattern is recognised as a grid
// and a tile's window and halo fit one workgroup's LDS; one workgroup per
// tile, at most one per CU, so every tile is resident.
void build_tiled(TrsvPlan &P, const std::vector<int> &iperm, const std::vector<int> &lrp,
                 hipStream_t s) {
  const int n = P.n;
  if (P.nnz == 0) {
    P.t_why = "no entries";
    return;
  }
  std::vector<int> level(static_cast<size_t>(n)), lpos(static_cast<size_t>(n));
  for (int l = 0; l < P.nlevels; ++l)
    for (int k = P.level_ptr[l]; k < P.level_ptr[l + 1]; ++k) {
      level[iperm[k]] = l;
      lpos[iperm[k]] = k;
    }
  int dev = 0, ncu = 0, budget = 0;
  LA_HIP(hipGetDevice(&dev));
  LA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  LA_HIP(hipDeviceGetAttribute(&budget, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  // halo slots: what LDS holds beyond the window, and what a 16-bit index
  // reaches
  const int64_t avail = static_cast<int64_t>(budget) - static_cast<int64_t>(tile_lds_bytes(0, 8));
  const int max_halo = static_cast<int>(std::min<int64_t>(avail / 8, 65535 - kTileXl - 1));
  if (max_halo < 128 || ncu < 1) {
    P.t_why = "LDS too small for the tiled schedule";
    return;
  }
  TileLayout L;
  if (!build_tile_layout(n, P.h_ai.data(), P.h_aj.data(), level.data(), kTileLanes, ncu, n,
                         max_halo, L, P.t_why))
    return;
  const int E = L.maxe, RL = L.rlanes;
  const int G = L.ngroups;
  const int64_t slots = static_cast<int64_t>(G) * kTileLanes, m = slots * E;
  std::vector<int> prow(static_cast<size_t>(slots), -1), dpos(static_cast<size_t>(slots), -1);
  std::vector<int> lsrc(static_cast<size_t>(m), -1);
  std::vector<uint16_t> pcode(static_cast<size_t>(m), static_cast<uint16_t>(kTileXl));
  std::vector<double> tv(static_cast<size_t>(m), 0.0), td(static_cast<size_t>(slots), 0.0);
  // a row of RL lanes: its first lane carries the row id, 1/d and the
  // level position; the others only entries (no id: they publish nothing)
  for (int g = 0; g < G; ++g)
    for (int kk = L.grp_row[g]; kk < L.grp_row[g + 1]; ++kk)
      for (int w = 0; w < RL; ++w) {
        const int64_t sl = static_cast<int64_t>(g) * kTileLanes + RL * (kk - L.grp_row[g]) + w;
        const int r = L.tperm[kk];
        if (w == 0) {
          prow[sl] = r;
          dpos[sl] = lpos[r];
          if (!P.unit) td[sl] = 1.0 / P.h_diag[r];
        }
        for (int e = 0; e < E; ++e) {
          const size_t te = (static_cast<size_t>(kk) * RL + w) * E + e;
          const size_t pe = static_cast<size_t>(sl) * E + e;
          const int c = L.tcode[te], q = L.tsrc[te];
          pcode[pe] = static_cast<uint16_t>(c >= 0 ? c : kTileXl + 1 + (-c - 1));
          lsrc[pe] = q >= 0 ? lrp[lpos[r]] + (q - P.h_ai[r]) : -1;
          tv[pe] = q >= 0 ? P.h_av[q] : 0.0;
        }
      }
  std::vector<int> rfslot(L.rf_slot.size());
  for (size_t t = 0; t < rfslot.size(); ++t) rfslot[t] = kTileXl + 1 + L.rf_slot[t];
  auto up = [&](DevBuf &d, const void *h, size_t bytes) {
    d.alloc(std::max<size_t>(bytes, 4));
    d.upload(h, bytes, s);
  };
  up(P.t_gptr, L.part_gptr.data(), sizeof(int) * L.part_gptr.size());
  up(P.t_prow, prow.data(), sizeof(int) * prow.size());
  up(P.t_pcode, pcode.data(), sizeof(uint16_t) * pcode.size());
  up(P.t_rfptr, L.rf_ptr.data(), sizeof(int) * L.rf_ptr.size());
  up(P.t_rfcol, L.rf_col.data(), sizeof(int) * L.rf_col.size());
  up(P.t_rfslot, rfslot.data(), sizeof(int) * rfslot.size());
  up(P.t_lsrc, lsrc.data(), sizeof(int) * lsrc.size());
  up(P.t_dpos, dpos.data(), sizeof(int) * dpos.size());
  P.t_val.set_from_host(tv.data(), m, s);
  if (!P.unit) P.t_dinv.set_from_host(td.data(), slots, s);
  P.t_nparts = L.nparts;
  P.t_ngroups = G;
  P.t_mmax = L.m_max;
  P.t_hmax = L.h_max;
  P.t_maxe = E;
  P.t_rlanes = RL;
  P.t_tau = L.tau;
  P.t_remote = L.remote_frac;
  P.t_wgs = static_cast<unsigned>(std::min(L.nparts, ncu));
  P.tiled = true;
}

hipStream_t capture_stream() {
  static thread_local hipStream_t cs = nullptr;
  if (!cs) LA_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  return cs;
}
}  // namespace

TrsvPlan::~TrsvPlan() {
  for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
}

std::unique_ptr<TrsvPlan> make_trsv(int n, int base, const int *ai, const int *aj,
                                    const double *av, const double *diag, bool upper,
                                    hipStream_t s) {
  LA_CHECK_ARG(n >= 0 && ai, "bad triangle");
  auto P = std::make_unique<TrsvPlan>();
  P->n = n;
  P->upper = upper;
  P->unit = diag == nullptr;
  P->stream = s;
  const int64_t nnz = static_cast<int64_t>(ai[n]) - ai[0];
  LA_CHECK_ARG(ai[0] == base, "ai[0] must equal base");
  LA_CHECK_ARG(nnz == 0 || (aj && av), "aj/av null");
  P->nnz = nnz;
  P->h_ai.resize(static_cast<size_t>(n) + 1);
  P->h_aj.resize(static_cast<size_t>(nnz));
  P->h_av.assign(av, av + nnz);
  for (int i = 0; i <= n; ++i) P->h_ai[i] = ai[i] - base;
  for (int i = 0; i < n; ++i)
    for (int q = P->h_ai[i]; q < P->h_ai[i + 1]; ++q) {
      const int c = aj[q] - base;
      LA_CHECK_ARG(c >= 0 && c < n, "column out of range");
      LA_CHECK_ARG(upper ? c > i : c < i,
                   "triangular solve needs a strictly triangular matrix (diagonal apart)");
      P->h_aj[q] = c;
    }
  if (diag) P->h_diag.assign(diag, diag + n);
  // level sets
  std::vector<int> iperm(std::max(n, 1)), prefix(static_cast<size_t>(n) + 2);
  const int L = topo_sort2(n, 0, P->h_ai.data(), P->h_aj.data(), upper, iperm.data(), prefix.data());
  P->nlevels = L;
  P->level_ptr.assign(prefix.begin(), prefix.begin() + L + 1);
  P->h_perm.assign(iperm.begin(), iperm.begin() + n);
  // level-ordered copy
  std::vector<int> rp(static_cast<size_t>(n) + 1), cl(static_cast<size_t>(nnz));
  std::vector<double> vl(static_cast<size_t>(nnz)), dg(diag ? n : 0);
  rp[0] = 0;
  for (int k = 0; k < n; ++k) {
    const int r = iperm[k];
    const int len = P->h_ai[r + 1] - P->h_ai[r];
    std::copy(P->h_aj.begin() + P->h_ai[r], P->h_aj.begin() + P->h_ai[r + 1], cl.begin() + rp[k]);
    std::copy(P->h_av.begin() + P->h_ai[r], P->h_av.begin() + P->h_ai[r + 1], vl.begin() + rp[k]);
    rp[k + 1] = rp[k] + len;
    if (diag) dg[k] = 1.0 / diag[r];  // kernels multiply: no divide on the chain
  }
  P->G = pick_G(n ? static_cast<double>(nnz) / n : 1.0);

  // sync-free launch shape. Narrow levels (few rows per level): the solve is
  // a chain of dependency hops, so keep it on one XCD, where a hop costs
  // ~0.3 us instead of ~0.55 us, with 32 workgroups of 256. Wide levels:
  // the whole chip, 256 workgroups of 512 (fewer tickets per level). The
  // 1024 rows/level switch point sits between the 512^2 5-point grid (256
  // rows/level, 6x faster on one XCD than level launches) and the RCM 128^3
  // 7-point grid (5490 rows/level, 3.3x faster over all XCDs).
  const double rows_per_level = L > 0 ? static_cast<double>(n) / L : 0.0;
  if (rows_per_level < 1024.0) {
    P->sf_xlocal = true;
    P->sf_wgs = 32;
    P->sf_nt = 256;
  } else {
    // wide levels: the whole chip, 256 workgroups; rows of 16+ lanes get
    // 1024-thread workgroups, so a level costs half the tickets (27-point
    // ILU(1) at 160^3: 5.4 ms per apply against 6.9 ms with 512)
    P->sf_xlocal = false;
    P->sf_wgs = 256;
    P->sf_nt = P->G >= 16 ? 1024 : 512;
  }
  // chunks: one workgroup's worth of rows (sf_nt/G), never spanning a level
  if (const char *e = std::getenv("LAAMD_SF_NT")) {
    const int v = std::atoi(e);
    P->sf_nt = v >= 1024 ? 1024 : v >= 512 ? 512 : 256;
  }
  const int R = P->sf_nt / P->G;
  std::vector<int> chunks;
  for (int l = 0; l < L; ++l)
    for (int k = P->level_ptr[l]; k < P->level_ptr[l + 1]; k += R) chunks.push_back(k);
  chunks.push_back(n);
  P->nchunks = static_cast<int>(chunks.size()) - 1;
  // launch-shape overrides for tuning: LAAMD_SF_XLOCAL=0|1, LAAMD_SF_WGS=1..4096
  if (const char *e = std::getenv("LAAMD_SF_XLOCAL")) P->sf_xlocal = std::atoi(e) != 0;
  if (const char *e = std::getenv("LAAMD_SF_WGS"))
    P->sf_wgs = static_cast<unsigned>(std::clamp(std::atoi(e), 1, 4096));
  P->rowptr.alloc(sizeof(int) * (static_cast<size_t>(n) + 1));
  P->col.alloc(sizeof(int) * static_cast<size_t>(nnz));
  P->perm.alloc(sizeof(int) * static_cast<size_t>(std::max(n, 1)));
  P->chunk_start.alloc(sizeof(int) * chunks.size());
  P->counter.alloc(sizeof(unsigned) * 16);
  P->err_flag.alloc(sizeof(int) * 4);
  P->rowptr.upload(rp.data(), sizeof(int) * (n + 1), s);
  P->col.upload(cl.data(), sizeof(int) * nnz, s);
  P->perm.upload(iperm.data(), sizeof(int) * n, s);
  P->chunk_start.upload(chunks.data(), sizeof(int) * chunks.size(), s);
  LA_HIP(hipMemsetAsync(P->err_flag.p, 0, sizeof(int) * 4, s));
  LA_HIP(hipMemsetAsync(P->counter.p, 0, sizeof(unsigned) * 16, s));
  P->val.set_from_host(vl.data(), nnz, s);
  if (diag) P->diag.set_from_host(dg.data(), n, s);
  const char *te = std::getenv("LAAMD_SF_TILED");
  if (!(te && te[0] == '0') && n > 0) build_tiled(*P, iperm, rp, s);
  LA_HIP(hipStreamSynchronize(s));
  return P;
}

void trsv_add_dtype(TrsvPlan &T, la_dtype dt) {
  if (T.nnz) T.val.add(dt, T.stream);
  else if (!T.val.has[LA_F64]) { /* empty triangle: nothing stored */ }
  if (!T.unit) T.diag.add(dt, T.stream);
  if (T.tiled) {
    if (T.nnz) T.t_val.add(dt, T.stream);
    if (!T.unit) T.t_dinv.add(dt, T.stream);
  }
  if (T.transposed) trsv_add_dtype(*T.transposed, dt);
  LA_HIP(hipStreamSynchronize(T.stream));
}

void trsv_solve(TrsvPlan &T, la_trsv_mode mode, la_dtype val, la_dtype vec, const void *b,
                void *x, hipStream_t s) {
  LA_CHECK_ARG(mode == LA_TRSV_LEVEL || mode == LA_TRSV_SYNCFREE || mode == LA_TRSV_JACOBI,
               "bad trsv mode");
  LA_CHECK_ARG(val >= LA_F64 && val <= LA_BF16, "bad value dtype");
  LA_CHECK_ARG(vec == LA_F64 || vec == LA_F32, "vectors must be FP64 or FP32");
  LA_CHECK_ARG(T.n == 0 || (b && x), "null vector");
  LA_CHECK_ARG(b != x || T.n == 0, "b and x must be distinct");
  if (T.n == 0) return;
  if (mode == LA_TRSV_JACOBI) {
    jacobi_solve(T, val, vec, b, x, s);
    return;
  }
  if ((T.nnz && !T.val.has[val]) || (!T.unit && !T.diag.has[val])) trsv_add_dtype(T, val);
  if (mode == LA_TRSV_SYNCFREE || T.nlevels <= 2) {
    solve_dispatch(T, mode, val, vec, b, x, s);
    return;
  }
  // LEVEL: replay a captured graph of the per-level launches
  auto key = std::make_tuple(static_cast<int>(val), static_cast<int>(vec), b, x);
  auto it = T.graphs.find(key);
  if (it == T.graphs.end()) {
    if (T.graphs.size() >= 16) {
      for (auto &kv : T.graphs) (void)hipGraphExecDestroy(kv.second);
      T.graphs.clear();
    }
    hipStream_t cs = capture_stream();
    hipGraph_t g = nullptr;
    LA_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    try {
      solve_dispatch(T, mode, val, vec, b, x, cs);
    } catch (...) {
      (void)hipStreamEndCapture(cs, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    LA_HIP(hipStreamEndCapture(cs, &g));
    hipGraphExec_t ge = nullptr;
    LA_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    it = T.graphs.emplace(key, ge).first;
  }
  LA_HIP(hipGraphLaunch(it->second, s));
}

TrsvPlan &trsv_transposed(TrsvPlan &T) {
  if (!T.transposed) {
    std::vector<int> ti(static_cast<size_t>(T.n) + 1), tj(static_cast<size_t>(T.nnz));
    std::vector<double> tv(static_cast<size_t>(T.nnz));
    transpose(T.n, T.n, 0, T.h_ai.data(), T.h_aj.data(), T.h_av.data(), ti.data(), tj.data(),
              tv.data());
    T.transposed = make_trsv(T.n, 0, ti.data(), tj.data(), tv.data(),
                             T.unit ? nullptr : T.h_diag.data(), !T.upper, T.stream);
    T.transposed->sweeps = T.sweeps;
    if (!T.h_src.empty() || !T.h_dsrc.empty()) {
      // the same transpose carries the source positions (exact as doubles)
      std::vector<double> sv(T.h_src.begin(), T.h_src.end()), st(static_cast<size_t>(T.nnz));
      std::vector<int> ti2(static_cast<size_t>(T.n) + 1), tj2(static_cast<size_t>(T.nnz));
      transpose(T.n, T.n, 0, T.h_ai.data(), T.h_aj.data(), sv.data(), ti2.data(), tj2.data(),
                st.data());
      std::vector<int64_t> tsrc(st.begin(), st.end());
      trsv_set_sources(*T.transposed, tsrc.data(), T.h_dsrc.empty() ? nullptr : T.h_dsrc.data(),
                       T.stream);
      if (T.fsrc) trsv_refresh(*T.transposed, T.fsrc, T.stream);
    }
    for (int dt = 1; dt < 3; ++dt)
      if (T.val.has[dt] || T.diag.has[dt]) trsv_add_dtype(*T.transposed, static_cast<la_dtype>(dt));
  }
  return *T.transposed;
}

namespace {
__global__ void k_gather(int64_t m, const int64_t *__restrict__ src, const double *__restrict__ f,
                         double *out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < m) out[t] = f[src[t]];
}
__global__ void k_gather_recip(int64_t m, const int64_t *__restrict__ src,
                               const double *__restrict__ f, double *out) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < m) out[t] = 1.0 / f[src[t]];
}
}  // namespace

void trsv_set_sources(TrsvPlan &T, const int64_t *src, const int64_t *dsrc, hipStream_t s) {
  const int n = T.n;
  T.jac.reset();  // the JACOBI copy is rebuilt with the new sources
  T.h_src.assign(src, src + T.nnz);
  if (dsrc) T.h_dsrc.assign(dsrc, dsrc + n);
  else T.h_dsrc.clear();
  // level order: entries of row perm[k] in their stored order
  std::vector<int64_t> ls(static_cast<size_t>(std::max<int64_t>(T.nnz, 1))),
      ld(static_cast<size_t>(std::max(n, 1)));
  int64_t p = 0;
  for (int k = 0; k < n; ++k) {
    const int r = T.h_perm[k];
    for (int q = T.h_ai[r]; q < T.h_ai[r + 1]; ++q) ls[p++] = src[q];
    if (dsrc) ld[k] = dsrc[r];
  }
  T.val_src.alloc(sizeof(int64_t) * ls.size());
  T.val_src.upload(ls.data(), sizeof(int64_t) * ls.size(), s);
  if (dsrc) {
    T.diag_src.alloc(sizeof(int64_t) * ld.size());
    T.diag_src.upload(ld.data(), sizeof(int64_t) * ld.size(), s);
  }
  LA_HIP(hipStreamSynchronize(s));
}

void trsv_refresh(TrsvPlan &T, const double *f, hipStream_t s) {
  T.fsrc = f;
  if (T.nnz) {
    double *v = T.val.v[LA_F64].as<double>();
    hipLaunchKernelGGL(k_gather, dim3(static_cast<unsigned>((T.nnz + 255) / 256)), dim3(256), 0,
                       s, T.nnz, T.val_src.as<int64_t>(), f, v);
    LA_HIP(hipGetLastError());
    T.val.set_from_device_f64(v, T.nnz, s);
  }
  if (!T.unit && !T.h_dsrc.empty() && T.n) {
    double *d = T.diag.v[LA_F64].as<double>();
    hipLaunchKernelGGL(k_gather_recip, dim3(static_cast<unsigned>((T.n + 255) / 256)), dim3(256),
                       0, s, static_cast<int64_t>(T.n), T.diag_src.as<int64_t>(), f, d);
    LA_HIP(hipGetLastError());
    T.diag.set_from_device_f64(d, T.n, s);
  }
  tiled_refresh(T, s);
  if (T.jac) jacobi_refresh(T, f, s);
  if (T.transposed) trsv_refresh(*T.transposed, f, s);
}

void trsv_set_sweeps(TrsvPlan &T, int sweeps) {
  LA_CHECK_ARG(sweeps >= 0 && sweeps <= 1000000, "sweeps out of range");
  T.sweeps = sweeps;
  if (T.transposed) trsv_set_sweeps(*T.transposed, sweeps);
}

bool trsv_check_timeout(TrsvPlan &T, hipStream_t s) {
  int flag = 0;
  LA_HIP(hipMemcpyAsync(&flag, T.err_flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  LA_HIP(hipStreamSynchronize(s));
  if (flag) LA_HIP(hipMemsetAsync(T.err_flag.p, 0, sizeof(int), s));
  bool bad = flag != 0;
  if (T.transposed) bad = trsv_check_timeout(*T.transposed, s) || bad;
  return bad;
}

// ------------------------------------------------------------------ ILU / IC
std::unique_ptr<Prec> make_prec_ilu(int n, int base, const int *fai, const int *faj,
                                    const double *fav, hipStream_t s) {
  auto P = std::make_unique<Prec>();
  P->n = n;
  P->stream = s;
  std::vector<int> Lai(static_cast<size_t>(n) + 1), Uai(static_cast<size_t>(n) + 1);
  int64_t ln = 0, un = 0;
  split_ldu(n, base, fai, faj, fav, Lai.data(), nullptr, nullptr, nullptr, Uai.data(), nullptr,
            nullptr, &ln, &un);
  std::vector<int> Laj(ln), Uaj(un);
  std::vector<double> Lav(ln), Uav(un), D(n);
  split_ldu(n, base, fai, faj, fav, Lai.data(), Laj.data(), Lav.data(), D.data(), Uai.data(),
            Uaj.data(), Uav.data(), &ln, &un);
  for (int i = 0; i < n; ++i)
    if (D[i] == 0.0) throw la_error(LA_ERR_NUMERIC, "ILU factor has a zero or missing diagonal");
  P->lower = make_trsv(n, base, Lai.data(), Laj.data(), Lav.data(), nullptr, false, s);
  P->upper = make_trsv(n, base, Uai.data(), Uaj.data(), Uav.data(), D.data(), true, s);
  P->tmp[0].alloc(sizeof(double) * std::max(n, 1));
  P->tmp[1].alloc(sizeof(float) * std::max(n, 1));
  return P;
}

std::unique_ptr<Prec> make_prec_ic(int n, int base, const int *rai, const int *raj,
                                   const double *rav, hipStream_t s) {
  auto P = std::make_unique<Prec>();
  P->ic = true;
  P->n = n;
  P->stream = s;
  std::vector<int> Lai(static_cast<size_t>(n) + 1), Uai(static_cast<size_t>(n) + 1);
  int64_t ln = 0, un = 0;
  split_ldu(n, base, rai, raj, rav, Lai.data(), nullptr, nullptr, nullptr, Uai.data(), nullptr,
            nullptr, &ln, &un);
  LA_CHECK_ARG(ln == 0, "IC factor must be upper triangular");
  std::vector<int> Uaj(un);
  std::vector<double> Uav(un), D(n);
  split_ldu(n, base, rai, raj, rav, Lai.data(), nullptr, nullptr, D.data(), Uai.data(),
            Uaj.data(), Uav.data(), &ln, &un);
  for (int i = 0; i < n; ++i)
    if (!(D[i] > 0.0)) throw la_error(LA_ERR_NUMERIC, "IC factor has a non-positive diagonal");
  std::vector<int> ti(static_cast<size_t>(n) + 1), tj(un);
  std::vector<double> tv(un);
  transpose(n, n, base, Uai.data(), Uaj.data(), Uav.data(), ti.data(), tj.data(), tv.data());
  P->lower = make_trsv(n, base, ti.data(), tj.data(), tv.data(), D.data(), false, s);
  P->upper = make_trsv(n, base, Uai.data(), Uaj.data(), Uav.data(), D.data(), true, s);
  P->tmp[0].alloc(sizeof(double) * std::max(n, 1));
  P->tmp[1].alloc(sizeof(float) * std::max(n, 1));
  return P;
}

void prec_add_dtype(Prec &P, la_dtype dt) {
  trsv_add_dtype(*P.lower, dt);
  trsv_add_dtype(*P.upper, dt);
}

void prec_apply(Prec &P, la_trsv_mode mode, la_dtype val, la_dtype vec, const void *r, void *z,
                hipStream_t s) {
  void *t = P.tmp[vec == LA_F64 ? 0 : 1].p;
  trsv_solve(*P.lower, mode, val, vec, r, t, s);
  trsv_solve(*P.upper, mode, val, vec, t, z, s);
}

void prec_apply_transpose(Prec &P, la_trsv_mode mode, la_dtype val, la_dtype vec,
                          const void *r, void *z, hipStream_t s) {
  if (P.ic) return prec_apply(P, mode, val, vec, r, z, s);  // M = R^T R is symmetric
  void *t = P.tmp[vec == LA_F64 ? 0 : 1].p;
  // (LU)^-T = L^-T U^-T: solve U^T first (a lower solve), then L^T
  trsv_solve(trsv_transposed(*P.upper), mode, val, vec, r, t, s);
  trsv_solve(trsv_transposed(*P.lower), mode, val, vec, t, z, s);
}

}  // namespace laamd
