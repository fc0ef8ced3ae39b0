// Vector kernels of the Krylov loops: FP64-accumulated dot products
// (deterministic two-stage reduction), fused CG updates with device-side
// scalars, multi-dot / multi-axpy for GMRES orthogonalisation, precision
// conversions.
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "dev_util.hpp"
#include "device.hpp"

namespace laamd {
using namespace dev;

namespace {
constexpr int kB = 256;

template <typename V>
__device__ __forceinline__ double block_reduce(double v, double *sh) {
  v = wave_sum(v);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kB / 64; ++w) t += sh[w];
  return t;  // valid in thread 0
}

template <typename V>
__global__ __launch_bounds__(kB) void k_dot_partial(int64_t n, const V *__restrict__ x,
                                                    const V *__restrict__ y, double *part) {
  __shared__ double sh[kB / 64];
  double acc = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kB)
    acc += static_cast<double>(x[i]) * static_cast<double>(y[i]);
  const double t = block_reduce<V>(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out[i] = sum_b part[i*nb + b], fixed order; one workgroup per i
__global__ void k_reduce_partials(int k, int nb, const double *__restrict__ part, double *out) {
  __shared__ double sh[kB / 64];
  for (int i = blockIdx.x; i < k; i += gridDim.x) {
    double acc = 0;
    for (int b = threadIdx.x; b < nb; b += kB) acc += part[static_cast<int64_t>(i) * nb + b];
    acc = wave_sum(acc);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0;
      for (int w = 0; w < kB / 64; ++w) t += sh[w];
      out[i] = t;
    }
    __syncthreads();
  }
}

template <typename V>
__global__ __launch_bounds__(kB) void k_multidot_partial(int64_t n, int k,
                                                         const V *__restrict__ Vm, int64_t ld,
                                                         const V *__restrict__ w,
                                                         double *part) {
  __shared__ double sh[kB / 64];
  // one pass per basis vector; w is re-read from L2
  for (int i = 0; i < k; ++i) {
    const V *v = Vm + static_cast<int64_t>(i) * ld;
    double acc = 0;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
         t += static_cast<int64_t>(gridDim.x) * kB)
      acc += static_cast<double>(v[t]) * static_cast<double>(w[t]);
    const double s = block_reduce<V>(acc, sh);
    if (threadIdx.x == 0) part[static_cast<int64_t>(i) * gridDim.x + blockIdx.x] = s;
    __syncthreads();
  }
}

template <typename V>
__global__ __launch_bounds__(kB) void k_multiaxpy(int64_t n, int k, const V *__restrict__ Vm,
                                                  int64_t ld, const double *__restrict__ h,
                                                  V *w) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB) {
    double acc = static_cast<double>(w[t]);
    for (int i = 0; i < k; ++i) acc -= h[i] * static_cast<double>(Vm[static_cast<int64_t>(i) * ld + t]);
    w[t] = static_cast<V>(acc);
  }
}

template <typename V>
__global__ void k_axpby(int64_t n, double a, const V *__restrict__ x, double b, V *y) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB) {
    const double yv = b == 0.0 ? 0.0 : b * static_cast<double>(y[t]);
    y[t] = static_cast<V>(a * static_cast<double>(x[t]) + yv);
  }
}

template <typename S, typename D>
__global__ void k_copy(int64_t n, const S *__restrict__ s, D *d, double scale) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB)
    d[t] = static_cast<D>(static_cast<double>(s[t]) * scale);
}

template <typename V>
__global__ void k_zero(int64_t n, V *x) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB)
    x[t] = V(0);
}

// Sum of the nb = kRedBlocks per-block partials of a dot product, in a fixed
// order, returned to every thread of the block. Every block computes the
// same value bit for bit, so a consumer kernel needs no separate reduction
// launch.
__device__ __forceinline__ double block_sum_partials(const double *__restrict__ part,
                                                     double *sh) {
  double v = 0;
  for (int b = threadIdx.x; b < kRedBlocks; b += kB) v += part[b];
  v = wave_sum(v);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0;
  for (int w = 0; w < kB / 64; ++w) t += sh[w];
  return t;
}

// ---- 16-byte vector forms of the CG kernels: each thread keeps four
// 16-byte loads per operand in flight (the scalar grid-stride forms leave
// these 2-16 MB streams latency-bound); used when every operand is 16-byte
// aligned. Element i's thread and summation order are fixed by n and the
// grid, so results stay run-to-run identical.
template <typename V> struct V16;
template <> struct V16<float> {
  using T = float4;
  static constexpr int k = 4;
};
template <> struct V16<double> {
  using T = double2;
  static constexpr int k = 2;
};
template <typename T, typename V>
__device__ __forceinline__ V &vel(T &a, int e) {
  return reinterpret_cast<V *>(&a)[e];
}
constexpr int kUn = 4;  // vectors in flight per operand and thread
inline bool al16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// <V_i, w> for i < k with 16-byte loads: the basis vectors are taken kMdG
// at a time, each thread reading w once per group and the group's vectors
// beside it (kMdG + 1 loads in flight), one FP64 accumulator per vector.
// Block b writes part[i * gridDim.x + b] (the layout k_reduce_partials
// sums); the order of every sum is fixed by n and the grid. Needs V, w and
// the stride ld 16-byte aligned.
constexpr int kMdG = 8;
template <typename V>
__global__ __launch_bounds__(kB) void k_multidot_partial_v(int64_t n, int k,
                                                           const V *__restrict__ Vm, int64_t ld,
                                                           const V *__restrict__ w,
                                                           double *part) {
  using T = typename V16<V>::T;
  constexpr int K = V16<V>::k;
  __shared__ double sh[kMdG][kB / 64];
  const T *wv = reinterpret_cast<const T *>(w);
  const int64_t nv = n / K, stride = static_cast<int64_t>(gridDim.x) * kB;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i0 = 0; i0 < k; i0 += kMdG) {
    const int g = k - i0 < kMdG ? k - i0 : kMdG;
    double acc[kMdG];
#pragma unroll
    for (int e = 0; e < kMdG; ++e) acc[e] = 0;
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; v < nv; v += stride) {
      T b = wv[v];
      T a[kMdG];
#pragma unroll
      for (int e = 0; e < kMdG; ++e)
        if (e < g) a[e] = reinterpret_cast<const T *>(Vm + static_cast<int64_t>(i0 + e) * ld)[v];
#pragma unroll
      for (int e = 0; e < kMdG; ++e)
        if (e < g)
#pragma unroll
          for (int c = 0; c < K; ++c)
            acc[e] += static_cast<double>(vel<T, V>(a[e], c)) *
                      static_cast<double>(vel<T, V>(b, c));
    }
    if (blockIdx.x == 0 && threadIdx.x < n - nv * K) {
      const int64_t t = nv * K + threadIdx.x;
      for (int e = 0; e < g; ++e)
        acc[e] += static_cast<double>(Vm[static_cast<int64_t>(i0 + e) * ld + t]) *
                  static_cast<double>(w[t]);
    }
#pragma unroll
    for (int e = 0; e < kMdG; ++e) {
      const double r = wave_sum(acc[e]);
      if (lane == 0) sh[e][wave] = r;
    }
    __syncthreads();
    if (threadIdx.x < g) {
      double t = 0;
      for (int q = 0; q < kB / 64; ++q) t += sh[threadIdx.x][q];
      part[static_cast<int64_t>(i0 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
    }
    __syncthreads();
  }
}

// w -= sum_i h_i V_i with 16-byte loads (same arithmetic per element as
// k_multiaxpy: FP64 accumulation in index order, one rounding to V). Needs
// V, w and the stride ld 16-byte aligned.
template <typename V>
__global__ __launch_bounds__(kB) void k_multiaxpy_v(int64_t n, int k, const V *__restrict__ Vm,
                                                    int64_t ld, const double *__restrict__ h,
                                                    V *w) {
  using T = typename V16<V>::T;
  constexpr int K = V16<V>::k;
  T *wv = reinterpret_cast<T *>(w);
  const int64_t nv = n / K, stride = static_cast<int64_t>(gridDim.x) * kB;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; v < nv; v += stride) {
    T b = wv[v];
    double acc[K];
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] = static_cast<double>(vel<T, V>(b, c));
    for (int i = 0; i < k; ++i) {
      T a = reinterpret_cast<const T *>(Vm + static_cast<int64_t>(i) * ld)[v];
      const double hi = h[i];
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] -= hi * static_cast<double>(vel<T, V>(a, c));
    }
#pragma unroll
    for (int c = 0; c < K; ++c) vel<T, V>(b, c) = static_cast<V>(acc[c]);
    wv[v] = b;
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * K) {
    const int64_t t = nv * K + threadIdx.x;
    double acc = static_cast<double>(w[t]);
    for (int i = 0; i < k; ++i) acc -= h[i] * static_cast<double>(Vm[static_cast<int64_t>(i) * ld + t]);
    w[t] = static_cast<V>(acc);
  }
}

template <typename V>
__global__ __launch_bounds__(kB) void k_dot_partial_v(int64_t n, const V *__restrict__ x,
                                                      const V *__restrict__ y, double *part) {
  using T = typename V16<V>::T;
  constexpr int K = V16<V>::k;
  __shared__ double sh[kB / 64];
  const T *xv = reinterpret_cast<const T *>(x), *yv = reinterpret_cast<const T *>(y);
  const int64_t nv = n / K, stride = static_cast<int64_t>(gridDim.x) * kB;
  int64_t v = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  double acc = 0;
  for (; v + (kUn - 1) * stride < nv; v += kUn * stride) {
    T a[kUn], b[kUn];
#pragma unroll
    for (int u = 0; u < kUn; ++u) {
      a[u] = xv[v + u * stride];
      b[u] = yv[v + u * stride];
    }
#pragma unroll
    for (int u = 0; u < kUn; ++u)
#pragma unroll
      for (int e = 0; e < K; ++e)
        acc += static_cast<double>(vel<T, V>(a[u], e)) * static_cast<double>(vel<T, V>(b[u], e));
  }
  for (; v < nv; v += stride) {
    T a = xv[v], b = yv[v];
#pragma unroll
    for (int e = 0; e < K; ++e)
      acc += static_cast<double>(vel<T, V>(a, e)) * static_cast<double>(vel<T, V>(b, e));
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * K) {
    const int64_t i = nv * K + threadIdx.x;
    acc += static_cast<double>(x[i]) * static_cast<double>(y[i]);
  }
  const double t = block_reduce<V>(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <typename V>
__global__ __launch_bounds__(kB) void k_cg_update_xr_v(int64_t n, V *x, V *r,
                                                       const V *__restrict__ p,
                                                       const V *__restrict__ q,
                                                       const double *__restrict__ part_rz,
                                                       const double *__restrict__ part_pq,
                                                       double *part_rr) {
  using T = typename V16<V>::T;
  constexpr int K = V16<V>::k;
  __shared__ double sh[kB / 64];
  const double rz = block_sum_partials(part_rz, sh);
  const double pq = block_sum_partials(part_pq, sh);
  const V a = static_cast<V>(rz / pq);
  T *xv = reinterpret_cast<T *>(x), *rv = reinterpret_cast<T *>(r);
  const T *pv = reinterpret_cast<const T *>(p), *qv = reinterpret_cast<const T *>(q);
  const int64_t nv = n / K, stride = static_cast<int64_t>(gridDim.x) * kB;
  int64_t v = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  double acc = 0;
  for (; v + (kUn - 1) * stride < nv; v += kUn * stride) {
    T xa[kUn], ra[kUn], pa[kUn], qa[kUn];
#pragma unroll
    for (int u = 0; u < kUn; ++u) {
      xa[u] = xv[v + u * stride];
      ra[u] = rv[v + u * stride];
      pa[u] = pv[v + u * stride];
      qa[u] = qv[v + u * stride];
    }
#pragma unroll
    for (int u = 0; u < kUn; ++u) {
#pragma unroll
      for (int e = 0; e < K; ++e) {
        vel<T, V>(xa[u], e) += a * vel<T, V>(pa[u], e);
        const V rn = vel<T, V>(ra[u], e) - a * vel<T, V>(qa[u], e);
        vel<T, V>(ra[u], e) = rn;
        acc += static_cast<double>(rn) * static_cast<double>(rn);
      }
      xv[v + u * stride] = xa[u];
      rv[v + u * stride] = ra[u];
    }
  }
  for (; v < nv; v += stride) {
    T xa = xv[v], ra = rv[v], pa = pv[v], qa = qv[v];
#pragma unroll
    for (int e = 0; e < K; ++e) {
      vel<T, V>(xa, e) += a * vel<T, V>(pa, e);
      const V rn = vel<T, V>(ra, e) - a * vel<T, V>(qa, e);
      vel<T, V>(ra, e) = rn;
      acc += static_cast<double>(rn) * static_cast<double>(rn);
    }
    xv[v] = xa;
    rv[v] = ra;
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * K) {
    const int64_t i = nv * K + threadIdx.x;
    x[i] += a * p[i];
    const V rn = r[i] - a * q[i];
    r[i] = rn;
    acc += static_cast<double>(rn) * static_cast<double>(rn);
  }
  __syncthreads();  // every wave is past its reads of sh in block_sum_partials
  const double s = block_reduce<V>(acc, sh);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = s;
}

template <typename V>
__global__ __launch_bounds__(kB) void k_cg_update_p_v(int64_t n, V *p, const V *__restrict__ z,
                                                      const double *__restrict__ part_new,
                                                      const double *__restrict__ part_old) {
  using T = typename V16<V>::T;
  constexpr int K = V16<V>::k;
  __shared__ double sh[kB / 64];
  const double rn = block_sum_partials(part_new, sh);
  const double ro = block_sum_partials(part_old, sh);
  const V b = static_cast<V>(rn / ro);
  T *pv = reinterpret_cast<T *>(p);
  const T *zv = reinterpret_cast<const T *>(z);
  const int64_t nv = n / K, stride = static_cast<int64_t>(gridDim.x) * kB;
  int64_t v = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
  for (; v + (kUn - 1) * stride < nv; v += kUn * stride) {
    T pa[kUn], za[kUn];
#pragma unroll
    for (int u = 0; u < kUn; ++u) {
      pa[u] = pv[v + u * stride];
      za[u] = zv[v + u * stride];
    }
#pragma unroll
    for (int u = 0; u < kUn; ++u) {
#pragma unroll
      for (int e = 0; e < K; ++e) vel<T, V>(pa[u], e) = vel<T, V>(za[u], e) + b * vel<T, V>(pa[u], e);
      pv[v + u * stride] = pa[u];
    }
  }
  for (; v < nv; v += stride) {
    T pa = pv[v], za = zv[v];
#pragma unroll
    for (int e = 0; e < K; ++e) vel<T, V>(pa, e) = vel<T, V>(za, e) + b * vel<T, V>(pa, e);
    pv[v] = pa;
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * K) {
    const int64_t i = nv * K + threadIdx.x;
    p[i] = z[i] + b * p[i];
  }
}

// x += alpha p ; r -= alpha q with alpha = <r,z> / <p,q> from their
// partials; writes the partials of <r,r>
template <typename V>
__global__ __launch_bounds__(kB) void k_cg_update_xr(int64_t n, V *x, V *r, const V *__restrict__ p,
                                                     const V *__restrict__ q,
                                                     const double *__restrict__ part_rz,
                                                     const double *__restrict__ part_pq,
                                                     double *part_rr) {
  __shared__ double sh[kB / 64];
  const double rz = block_sum_partials(part_rz, sh);
  const double pq = block_sum_partials(part_pq, sh);
  const V a = static_cast<V>(rz / pq);
  double acc = 0;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB) {
    x[t] += a * p[t];
    const V rn = r[t] - a * q[t];
    r[t] = rn;
    acc += static_cast<double>(rn) * static_cast<double>(rn);
  }
  __syncthreads();  // every wave is past its reads of sh in block_sum_partials
  const double s = block_reduce<V>(acc, sh);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = s;
}

// p = z + beta p with beta = <r,z>_new / <r,z>_old from their partials
template <typename V>
__global__ __launch_bounds__(kB) void k_cg_update_p(int64_t n, V *p, const V *__restrict__ z,
                                                    const double *__restrict__ part_new,
                                                    const double *__restrict__ part_old) {
  __shared__ double sh[kB / 64];
  const double rn = block_sum_partials(part_new, sh);
  const double ro = block_sum_partials(part_old, sh);
  const V b = static_cast<V>(rn / ro);
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB)
    p[t] = z[t] + b * p[t];
}

template <typename V>
__global__ void k_add_scaled(int64_t n, const V *__restrict__ d, double scale, double *x) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB)
    x[t] += scale * static_cast<double>(d[t]);
}

template <typename V>
__global__ void k_scale_by_inv(int64_t n, const V *__restrict__ w, const double *__restrict__ h,
                               V *v) {
  const double inv = 1.0 / h[0];
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB)
    v[t] = static_cast<V>(static_cast<double>(w[t]) * inv);
}

// FP64 -> FP32 / BF16 (round to nearest even) on the device
__global__ void k_f64_to_f32(int64_t n, const double *__restrict__ s, float *d) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB)
    d[t] = static_cast<float>(s[t]);
}
__device__ __forceinline__ uint16_t dev_f64_to_bf16(double v) {
  // round the double directly to BF16 (8 significant bits), ties to even,
  // with no intermediate rounding to FP32
  unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
  const uint16_t sign = static_cast<uint16_t>((u >> 48) & 0x8000);
  const unsigned long long exp = (u >> 52) & 0x7FF;
  if (exp == 0x7FF) return sign | ((u << 12) ? 0x7FC0 : 0x7F80);  // nan / inf
  if (exp < 1023 - 126) {
    // below the smallest normal: BF16's subnormal grid is k * 2^-133; the
    // scaling is exact and rint rounds half to even (k = 128 is 2^-126)
    const double k = __builtin_rint(fabs(v) * 0x1p133);
    return sign | static_cast<uint16_t>(k);
  }
  if (exp > 1023 + 127) return sign | 0x7F80;  // |v| >= 2^128 overflows
  // add half an ulp of bit 45 with ties to even, truncate, narrow exactly
  u += (1ULL << 44) - 1 + ((u >> 45) & 1);
  u &= ~((1ULL << 45) - 1);
  const uint32_t f = __float_as_uint(static_cast<float>(__longlong_as_double(static_cast<long long>(u))));
  return static_cast<uint16_t>(f >> 16);
}
__global__ void k_f64_to_bf16(int64_t n, const double *__restrict__ s, bf16 *d) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * kB)
    d[t].bits = dev_f64_to_bf16(s[t]);
}

inline unsigned grid_for(int64_t n, int64_t cap = 4096) {
  int64_t g = (n + kB - 1) / kB;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}
}  // namespace

// ------------------------------------------------------------------ host
void launch_convert_f64(const double *src, int64_t n, la_dtype dst_dt, void *dst, hipStream_t s) {
  if (n == 0) return;
  if (dst_dt == LA_F32)
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_for(n)), dim3(kB), 0, s, n, src, static_cast<float *>(dst));
  else if (dst_dt == LA_BF16)
    hipLaunchKernelGGL(k_f64_to_bf16, dim3(grid_for(n)), dim3(kB), 0, s, n, src, static_cast<bf16 *>(dst));
  else
    LA_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  LA_HIP(hipGetLastError());
}

namespace {
void alloc_padded(DevBuf &b, la_dtype dt, int64_t cnt, hipStream_t s) {
  const int64_t pc = padded_count(cnt);
  b.alloc(dtype_size(dt) * static_cast<size_t>(pc > 0 ? pc : 1));
  if (pc > cnt)
    LA_HIP(hipMemsetAsync(static_cast<char *>(b.p) + dtype_size(dt) * cnt, 0,
                          dtype_size(dt) * static_cast<size_t>(pc - cnt), s));
}
}  // namespace

void MultiValues::set_from_device_f64(const double *d, int64_t cnt, hipStream_t s) {
  count = cnt;
  ++version;
  if (!has[LA_F64]) {
    alloc_padded(v[LA_F64], LA_F64, cnt, s);
    has[LA_F64] = true;
  }
  if (cnt && d != v[LA_F64].p)
    LA_HIP(hipMemcpyAsync(v[LA_F64].p, d, sizeof(double) * cnt, hipMemcpyDeviceToDevice, s));
  for (int dt = 1; dt < 3; ++dt)
    if (has[dt]) launch_convert_f64(v[LA_F64].as<double>(), cnt, static_cast<la_dtype>(dt), v[dt].p, s);
}

void MultiValues::set_from_host(const double *h, int64_t cnt, hipStream_t s) {
  count = cnt;
  ++version;
  if (!has[LA_F64]) {
    alloc_padded(v[LA_F64], LA_F64, cnt, s);
    has[LA_F64] = true;
  }
  v[LA_F64].upload(h, sizeof(double) * cnt, s);
  for (int dt = 1; dt < 3; ++dt)
    if (has[dt]) launch_convert_f64(v[LA_F64].as<double>(), cnt, static_cast<la_dtype>(dt), v[dt].p, s);
}

void MultiValues::add(la_dtype dt, hipStream_t s) {
  if (has[dt]) return;
  if (!has[LA_F64]) throw la_error(LA_ERR_STATE, "no FP64 values to convert from");
  alloc_padded(v[dt], dt, count, s);
  launch_convert_f64(v[LA_F64].as<double>(), count, dt, v[dt].p, s);
  has[dt] = true;
}

Scratch::Scratch() {
  partials.alloc(sizeof(double) * kRedBlocks * kMaxDots);
  scalars.alloc(sizeof(double) * 256);
  LA_HIP(hipHostMalloc(reinterpret_cast<void **>(&host), sizeof(double) * 2 * kRedBlocks,
                       hipHostMallocDefault));
  LA_HIP(hipHostMalloc(reinterpret_cast<void **>(&rep), sizeof(uint64_t) * 32,
                       hipHostMallocMapped | hipHostMallocCoherent));
  rep[0] = rep[1] = 0;
  LA_HIP(hipHostGetDevicePointer(reinterpret_cast<void **>(&rep_dev), rep, 0));
}
Scratch::~Scratch() {
  if (host) (void)hipHostFree(host);
  if (rep) (void)hipHostFree(rep);
}

namespace {
// the slot's "no report yet" value: a signalling-NaN pattern that no sum
// produces (arithmetic yields the canonical quiet NaN)
constexpr uint64_t kRepEmpty = 0x7FF4DEADBEEF0001ULL;

// one workgroup: the fixed-order sum of the partials (block_sum_partials'
// order), then its bits into host memory with one system-scope store (a
// single word carries the whole report, so no ordering is needed)
__global__ __launch_bounds__(kB) void k_report(const double *__restrict__ part, uint64_t *slot) {
  __shared__ double sh[kB / 64];
  const double t = block_sum_partials(part, sh);
  if (threadIdx.x == 0) {
    uint64_t bits = static_cast<uint64_t>(__double_as_longlong(t));
    if (bits == kRepEmpty) bits ^= 1;  // keep the marker unambiguous
    __hip_atomic_store(slot, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
}  // namespace

uint64_t report_sum(Scratch &S, const double *part, hipStream_t s) {
  // the slot is marked empty before the launch, which is safe only once the
  // previous report's kernel has landed: a report that was never read (an
  // error path left it) is drained first
  if (S.rep_pending) LA_HIP(hipStreamSynchronize(s));
  const uint64_t seq = ++S.rep_seq;
  S.rep_pending = seq;
  __atomic_store_n(S.rep, kRepEmpty, __ATOMIC_SEQ_CST);
  hipLaunchKernelGGL(k_report, dim3(1), dim3(kB), 0, s, part, S.rep_dev);
  LA_HIP(hipGetLastError());
  return seq;
}

double wait_report(Scratch &S, uint64_t seq, hipStream_t s) {
  // spin on the coherent slot; after ~2 s without it, let the runtime
  // report a stream error (or the value, if it came meanwhile)
  if (seq == 0 || seq != S.rep_pending)
    throw la_error(LA_ERR_STATE, "wait_report: not the outstanding report");
  S.rep_pending = 0;
  const auto t0 = std::chrono::steady_clock::now();
  uint64_t bits;
  for (uint64_t spins = 0;; ++spins) {
    if ((bits = __atomic_load_n(S.rep, __ATOMIC_ACQUIRE)) != kRepEmpty) break;
    if ((spins & 4095) == 4095 &&
        std::chrono::steady_clock::now() - t0 > std::chrono::seconds(2)) {
      LA_HIP(hipStreamSynchronize(s));
      if ((bits = __atomic_load_n(S.rep, __ATOMIC_ACQUIRE)) != kRepEmpty) break;
      throw la_error(LA_ERR_HIP, "device report did not arrive");
    }
  }
  double v;
  std::memcpy(&v, &bits, sizeof v);
  return v;
}

// ---------------------------------------------------- allocation-free dot
// The partial sums of dot_noalloc live in a fixed device array, one copy per
// GPU for the life of the process: kRedBlocks doubles (4 KiB), the grid of
// the reduction, whatever n is. Host code never passes `scratch` itself (in
// host code the name is only a stand-in); it uses the per-device address
// hipGetSymbolAddress returns, looked up once per device.
constexpr int kNoAllocScratch = kRedBlocks;
static __device__ __align__(256) double scratch[kNoAllocScratch];

namespace {
double *scratch_ptr() {
  static std::mutex mu;
  static std::array<double *, 64> ptrs{};
  int dev = 0;
  LA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= static_cast<int>(ptrs.size()))
    throw la_error(LA_ERR_ARG, "device index out of range for the fixed scratch table");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void *p = nullptr;
    LA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw la_error(LA_ERR_HIP, "no device address for the fixed scratch");
    ptrs[dev] = static_cast<double *>(p);
  }
  return ptrs[dev];
}
}  // namespace

void dot_noalloc(int64_t n, la_dtype vec, const void *x, const void *y, double *out_dev,
                 hipStream_t s) {
  double *part = scratch_ptr();
  static_assert(kNoAllocScratch >= kRedBlocks, "scratch holds one partial per block");
  // every block writes its partial (0 when it has no elements), so nothing
  // left by an earlier call is read: no clearing is needed
  dot_partials(n, vec, x, y, part, s);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kB), 0, s, 1, kRedBlocks, part, out_dev);
  LA_HIP(hipGetLastError());
}

void dot_dev(int64_t n, la_dtype vec, const void *x, const void *y, double *partials, double *out,
             hipStream_t s) {
  dot_partials(n, vec, x, y, partials, s);  // 16-byte loads when x and y allow them
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kB), 0, s, 1, kRedBlocks, partials, out);
  LA_HIP(hipGetLastError());
}

void multidot_dev(int64_t n, int k, la_dtype vec, const void *V, int64_t ld, const void *w,
                  double *partials, double *out, hipStream_t s) {
  LA_CHECK_ARG(k <= kMaxDots, "too many dot products");
  const size_t es = vec == LA_F64 ? sizeof(double) : sizeof(float);
  const bool v16 = al16(V) && al16(w) && (static_cast<size_t>(ld) * es) % 16 == 0;
  if (vec == LA_F64 && v16)
    hipLaunchKernelGGL(k_multidot_partial_v<double>, dim3(kRedBlocks), dim3(kB), 0, s, n, k,
                       static_cast<const double *>(V), ld, static_cast<const double *>(w), partials);
  else if (v16)
    hipLaunchKernelGGL(k_multidot_partial_v<float>, dim3(kRedBlocks), dim3(kB), 0, s, n, k,
                       static_cast<const float *>(V), ld, static_cast<const float *>(w), partials);
  else if (vec == LA_F64)
    hipLaunchKernelGGL(k_multidot_partial<double>, dim3(kRedBlocks), dim3(kB), 0, s, n, k,
                       static_cast<const double *>(V), ld, static_cast<const double *>(w), partials);
  else
    hipLaunchKernelGGL(k_multidot_partial<float>, dim3(kRedBlocks), dim3(kB), 0, s, n, k,
                       static_cast<const float *>(V), ld, static_cast<const float *>(w), partials);
  LA_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_partials, dim3(std::max(k, 1)), dim3(kB), 0, s, k, kRedBlocks,
                     partials, out);
  LA_HIP(hipGetLastError());
}

void multiaxpy_dev(int64_t n, int k, la_dtype vec, const void *V, int64_t ld, const double *h,
                   void *w, hipStream_t s) {
  const size_t es = vec == LA_F64 ? sizeof(double) : sizeof(float);
  if (al16(V) && al16(w) && (static_cast<size_t>(ld) * es) % 16 == 0) {
    if (vec == LA_F64)
      hipLaunchKernelGGL(k_multiaxpy_v<double>, dim3(grid_for(n / 2)), dim3(kB), 0, s, n, k,
                         static_cast<const double *>(V), ld, h, static_cast<double *>(w));
    else
      hipLaunchKernelGGL(k_multiaxpy_v<float>, dim3(grid_for(n / 4)), dim3(kB), 0, s, n, k,
                         static_cast<const float *>(V), ld, h, static_cast<float *>(w));
    LA_HIP(hipGetLastError());
    return;
  }
  if (vec == LA_F64)
    hipLaunchKernelGGL(k_multiaxpy<double>, dim3(grid_for(n)), dim3(kB), 0, s, n, k,
                       static_cast<const double *>(V), ld, h, static_cast<double *>(w));
  else
    hipLaunchKernelGGL(k_multiaxpy<float>, dim3(grid_for(n)), dim3(kB), 0, s, n, k,
                       static_cast<const float *>(V), ld, h, static_cast<float *>(w));
  LA_HIP(hipGetLastError());
}

void axpby(int64_t n, la_dtype vec, double a, const void *x, double b, void *y, hipStream_t s) {
  if (vec == LA_F64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(grid_for(n)), dim3(kB), 0, s, n, a,
                       static_cast<const double *>(x), b, static_cast<double *>(y));
  else
    hipLaunchKernelGGL(k_axpby<float>, dim3(grid_for(n)), dim3(kB), 0, s, n, a,
                       static_cast<const float *>(x), b, static_cast<float *>(y));
  LA_HIP(hipGetLastError());
}

void copy_vec(int64_t n, la_dtype sd, const void *src, la_dtype dd, void *dst, double scale,
              hipStream_t s) {
  const unsigned g = grid_for(n);
  if (sd == LA_F64 && dd == LA_F64)
    hipLaunchKernelGGL((k_copy<double, double>), dim3(g), dim3(kB), 0, s, n, static_cast<const double *>(src), static_cast<double *>(dst), scale);
  else if (sd == LA_F64 && dd == LA_F32)
    hipLaunchKernelGGL((k_copy<double, float>), dim3(g), dim3(kB), 0, s, n, static_cast<const double *>(src), static_cast<float *>(dst), scale);
  else if (sd == LA_F32 && dd == LA_F64)
    hipLaunchKernelGGL((k_copy<float, double>), dim3(g), dim3(kB), 0, s, n, static_cast<const float *>(src), static_cast<double *>(dst), scale);
  else
    hipLaunchKernelGGL((k_copy<float, float>), dim3(g), dim3(kB), 0, s, n, static_cast<const float *>(src), static_cast<float *>(dst), scale);
  LA_HIP(hipGetLastError());
}

void fill_zero(int64_t n, la_dtype vec, void *x, hipStream_t s) {
  if (vec == LA_F64)
    hipLaunchKernelGGL(k_zero<double>, dim3(grid_for(n)), dim3(kB), 0, s, n, static_cast<double *>(x));
  else
    hipLaunchKernelGGL(k_zero<float>, dim3(grid_for(n)), dim3(kB), 0, s, n, static_cast<float *>(x));
  LA_HIP(hipGetLastError());
}

void dot_partials(int64_t n, la_dtype vec, const void *x, const void *y, double *partials,
                  hipStream_t s) {
  if (al16(x) && al16(y)) {
    if (vec == LA_F64)
      hipLaunchKernelGGL(k_dot_partial_v<double>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                         static_cast<const double *>(x), static_cast<const double *>(y),
                         partials);
    else
      hipLaunchKernelGGL(k_dot_partial_v<float>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                         static_cast<const float *>(x), static_cast<const float *>(y), partials);
    LA_HIP(hipGetLastError());
    return;
  }
  if (vec == LA_F64)
    hipLaunchKernelGGL(k_dot_partial<double>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                       static_cast<const double *>(x), static_cast<const double *>(y), partials);
  else
    hipLaunchKernelGGL(k_dot_partial<float>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                       static_cast<const float *>(x), static_cast<const float *>(y), partials);
  LA_HIP(hipGetLastError());
}

void cg_update_xr(int64_t n, la_dtype vec, void *x, void *r, const void *p, const void *q,
                  const double *part_rz, const double *part_pq, double *part_rr, hipStream_t s) {
  if (al16(x) && al16(r) && al16(p) && al16(q)) {
    if (vec == LA_F64)
      hipLaunchKernelGGL(k_cg_update_xr_v<double>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                         static_cast<double *>(x), static_cast<double *>(r),
                         static_cast<const double *>(p), static_cast<const double *>(q),
                         part_rz, part_pq, part_rr);
    else
      hipLaunchKernelGGL(k_cg_update_xr_v<float>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                         static_cast<float *>(x), static_cast<float *>(r),
                         static_cast<const float *>(p), static_cast<const float *>(q), part_rz,
                         part_pq, part_rr);
    LA_HIP(hipGetLastError());
    return;
  }
  if (vec == LA_F64)
    hipLaunchKernelGGL(k_cg_update_xr<double>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                       static_cast<double *>(x), static_cast<double *>(r),
                       static_cast<const double *>(p), static_cast<const double *>(q), part_rz,
                       part_pq, part_rr);
  else
    hipLaunchKernelGGL(k_cg_update_xr<float>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                       static_cast<float *>(x), static_cast<float *>(r),
                       static_cast<const float *>(p), static_cast<const float *>(q), part_rz,
                       part_pq, part_rr);
  LA_HIP(hipGetLastError());
}

void cg_update_p(int64_t n, la_dtype vec, void *p, const void *z, const double *part_new,
                 const double *part_old, hipStream_t s) {
  if (al16(p) && al16(z)) {
    if (vec == LA_F64)
      hipLaunchKernelGGL(k_cg_update_p_v<double>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                         static_cast<double *>(p), static_cast<const double *>(z), part_new,
                         part_old);
    else
      hipLaunchKernelGGL(k_cg_update_p_v<float>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                         static_cast<float *>(p), static_cast<const float *>(z), part_new,
                         part_old);
    LA_HIP(hipGetLastError());
    return;
  }
  if (vec == LA_F64)
    hipLaunchKernelGGL(k_cg_update_p<double>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                       static_cast<double *>(p), static_cast<const double *>(z), part_new, part_old);
  else
    hipLaunchKernelGGL(k_cg_update_p<float>, dim3(kRedBlocks), dim3(kB), 0, s, n,
                       static_cast<float *>(p), static_cast<const float *>(z), part_new, part_old);
  LA_HIP(hipGetLastError());
}

void add_scaled(int64_t n, la_dtype dvec, const void *d, double scale, double *x64, hipStream_t s) {
  if (dvec == LA_F64)
    hipLaunchKernelGGL(k_add_scaled<double>, dim3(grid_for(n)), dim3(kB), 0, s, n, static_cast<const double *>(d), scale, x64);
  else
    hipLaunchKernelGGL(k_add_scaled<float>, dim3(grid_for(n)), dim3(kB), 0, s, n, static_cast<const float *>(d), scale, x64);
  LA_HIP(hipGetLastError());
}

void scale_by_inv(int64_t n, la_dtype vec, const void *w, const double *h, void *v, hipStream_t s) {
  if (vec == LA_F64)
    hipLaunchKernelGGL(k_scale_by_inv<double>, dim3(grid_for(n)), dim3(kB), 0, s, n, static_cast<const double *>(w), h, static_cast<double *>(v));
  else
    hipLaunchKernelGGL(k_scale_by_inv<float>, dim3(grid_for(n)), dim3(kB), 0, s, n, static_cast<const float *>(w), h, static_cast<float *>(v));
  LA_HIP(hipGetLastError());
}

}  // namespace laamd
