// Device helpers shared by the HIP kernels: typed loads (FP64/FP32/BF16
// storage), DPP group reductions over 64-wide wavefronts, readiness
// sentinels for the sync-free triangular solve.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace laamd {
namespace dev {

// ---------------------------------------------------------------- loads
// Matrix data is streamed once per product: non-temporal loads keep it from
// displacing the gathered vector in the L2.
__device__ __forceinline__ double ldm(const double *p, int64_t i) {
  return __builtin_nontemporal_load(p + i);
}
__device__ __forceinline__ float ldm(const float *p, int64_t i) {
  return __builtin_nontemporal_load(p + i);
}
__device__ __forceinline__ float ldm(const bf16 *p, int64_t i) {
  const uint16_t b = __builtin_nontemporal_load(&p[i].bits);
  return __uint_as_float(static_cast<uint32_t>(b) << 16);
}
__device__ __forceinline__ int ldi(const int *p, int64_t i) {
  return __builtin_nontemporal_load(p + i);
}

// ---------------------------------------------------------------- DPP
// dpp_ctrl encodings (GCN/CDNA DPP): quad_perm [1,0,3,2] = 0xB1 (xor 1),
// quad_perm [2,3,0,1] = 0x4E (xor 2), row_half_mirror = 0x141 (lane i <->
// 7-i inside 8 lanes), row_mirror = 0x140 (i <-> 15-i inside 16 lanes).
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, 0xF, 0xF, false);
  const int hi =
      __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}

// Sum over aligned groups of G lanes (G a power of two <= 64); every lane of
// the group gets the total. Lanes of one group must all be active.
template <int G, typename A>
__device__ __forceinline__ A group_sum(A v) {
  if constexpr (G >= 2) v += dpp<0xB1>(v);
  if constexpr (G >= 4) v += dpp<0x4E>(v);
  if constexpr (G >= 8) v += dpp<0x141>(v);
  if constexpr (G >= 16) v += dpp<0x140>(v);
  if constexpr (G >= 32) v += __shfl_xor(v, 16, 64);
  if constexpr (G >= 64) v += __shfl_xor(v, 32, 64);
  return v;
}

template <typename A>
__device__ __forceinline__ A wave_sum(A v) {
  return group_sum<64>(v);
}

// ---------------------------------------------------------------- sentinels
// A value that arithmetic never produces (a signalling-NaN payload) marks an
// x entry as not yet solved. A computed result that happens to carry the
// same bits is replaced by the canonical quiet NaN before it is published.
template <typename V> struct Sentinel;
template <> struct Sentinel<double> {
  using U = unsigned long long;
  static constexpr U bits = 0x7FF4A5A5A5A5A5A5ULL;
  static constexpr U qnan = 0x7FF8000000000000ULL;
};
template <> struct Sentinel<float> {
  using U = unsigned int;
  static constexpr U bits = 0x7FA5A5A5U;
  static constexpr U qnan = 0x7FC00000U;
};

template <typename V>
__device__ __forceinline__ typename Sentinel<V>::U to_bits(V v) {
  if constexpr (sizeof(V) == 8) return static_cast<unsigned long long>(__double_as_longlong(v));
  else return __float_as_uint(v);
}
template <typename V>
__device__ __forceinline__ V from_bits(typename Sentinel<V>::U u) {
  if constexpr (sizeof(V) == 8) return __longlong_as_double(static_cast<long long>(u));
  else return __uint_as_float(u);
}

// ---------------------------------------------------------------- XCD id
// MI355X has 8 XCDs, each with its own L2; workgroups are dealt to them
// round-robin in dispatch order.
// Block b of a launch of nb blocks, renumbered so that each XCD gets one
// contiguous range of logical blocks (a bijection on [0, nb) whatever the
// placement; with round-robin placement neighbouring work, and the vector
// lines it gathers, stays in one XCD's L2).
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
  const unsigned xcd = b & 7u, idx = b >> 3, per = nb >> 3, rem = nb & 7u;
  return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// the XCD this wavefront runs on (HW_REG_XCC_ID, bits 3:0)
__device__ __forceinline__ unsigned xcc_id() {
  unsigned v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return v & 0xF;
}

// Agent-scope relaxed atomics: one location carries both the value and its
// readiness, so no ordering against other locations is needed.
template <typename V>
__device__ __forceinline__ V load_ready(const V *p) {
  using U = typename Sentinel<V>::U;
  U u = __hip_atomic_load(reinterpret_cast<const U *>(p), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
  return from_bits<V>(u);
}
template <typename V>
__device__ __forceinline__ bool is_sentinel(V v) {
  return to_bits<V>(v) == Sentinel<V>::bits;
}
template <typename V>
__device__ __forceinline__ void publish(V *p, V v) {
  using U = typename Sentinel<V>::U;
  U u = to_bits<V>(v);
  if (u == Sentinel<V>::bits) u = Sentinel<V>::qnan;
  __hip_atomic_store(reinterpret_cast<U *>(p), u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace dev
}  // namespace laamd
