// Shared definitions for the laamd library: status handling, dtype traits,
// BF16 storage, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "laamd.h"

namespace laamd {

// ---------------------------------------------------------------- errors
// Internal code throws la_error; the C ABI layer catches everything and
// turns it into a status code plus a thread-local message.
struct la_error : std::runtime_error {
  la_status code;
  la_error(la_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define LA_CHECK_ARG(cond, msg)                                              \
  do {                                                                       \
    if (!(cond)) throw ::laamd::la_error(LA_ERR_ARG, (msg));                 \
  } while (0)

#define LA_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e__ = (call);                                                 \
    if (e__ != hipSuccess)                                                   \
      throw ::laamd::la_error(                                               \
          LA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e__));   \
  } while (0)

// ---------------------------------------------------------------- BF16
// BF16 values are stored as raw uint16_t (the upper half of an IEEE float)
// and widened with a shift; no library type is involved.
struct bf16 {
  uint16_t bits;
};


inline size_t dtype_size(la_dtype dt) {
  return dt == LA_F64 ? 8 : dt == LA_F32 ? 4 : 2;
}

// ---------------------------------------------------------------- device mem
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b == 0) b = 16;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      throw la_error(LA_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(b) +
                                       "): " + hipGetErrorString(e));
    }
    bytes = b;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
  void upload(const void *src, size_t b, hipStream_t s) {
    if (b) LA_HIP(hipMemcpyAsync(p, src, b, hipMemcpyHostToDevice, s));
  }
};

// Arrays streamed by the nnz-balanced SpMV are allocated to a whole number
// of its 512-entry segments and zero-filled past the end, so the kernel's
// vector loads never need a bounds branch.
constexpr int64_t kNnzPad = 512;
inline int64_t padded_count(int64_t n) { return ((n + kNnzPad - 1) / kNnzPad) * kNnzPad; }

// Values of a CSR (or factor) kept in up to three precisions (allocations
// padded, see padded_count).
struct MultiValues {
  DevBuf v[3];  // indexed by la_dtype
  bool has[3] = {false, false, false};
  int64_t count = 0;
  uint64_t version = 0;  // bumped whenever the values change
  // uploads FP64 host values and converts every requested precision
  void set_from_host(const double *h, int64_t cnt, hipStream_t s);
  void set_from_device_f64(const double *d, int64_t cnt, hipStream_t s);
  void add(la_dtype dt, hipStream_t s);
  const void *get(la_dtype dt) const {
    if (!has[dt]) throw la_error(LA_ERR_STATE, "values not stored in requested precision");
    return v[dt].p;
  }
};

// device-side conversion kernels (blas1.hip)
void launch_convert_f64(const double *src, int64_t n, la_dtype dst_dt, void *dst,
                        hipStream_t s);

}  // namespace laamd
