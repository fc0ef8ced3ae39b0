// Symmetric row-norm scaling of a device matrix (mkl_wrapper/
// mkl_sparse_mat.cpp:517-544, rowwiseSqrtNorm and DtAD), one thread per row
// in the host's order with every product rounded on its own: bitwise the
// host's results.
#include <cfloat>

#include "dev_util.hpp"
#include "device.hpp"

#pragma clang fp contract(off)

namespace laamd {
namespace {
__global__ void k_row_sqrt_norm(int n, const int *__restrict__ ai, const double *__restrict__ av,
                                double *d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = ai[i]; j < ai[i + 1]; ++j) s += av[j] * av[j];
  d[i] = s > DBL_MIN ? 1.0 / sqrt(sqrt(s)) : 1.0;
}
__global__ void k_dtad(int n, const int *__restrict__ ai, const int *__restrict__ aj, double *av,
                       const double *__restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double di = d[i];
  for (int j = ai[i]; j < ai[i + 1]; ++j) av[j] *= di * d[aj[j]];
}
}  // namespace

void mat_rowwise_sqrt_norm(Mat &A, double *d, hipStream_t s) {
  if (A.rows == 0) return;
  hipLaunchKernelGGL(k_row_sqrt_norm, dim3((A.rows + 255) / 256), dim3(256), 0, s, A.rows,
                     A.ai.as<int>(), static_cast<const double *>(A.vals.get(LA_F64)), d);
  LA_HIP(hipGetLastError());
}

void mat_dtad(Mat &A, const double *d, hipStream_t s) {
  LA_CHECK_ARG(A.rows == A.cols, "DtAD needs a square matrix");
  if (A.rows == 0) return;
  double *v = A.vals.v[LA_F64].as<double>();
  hipLaunchKernelGGL(k_dtad, dim3((A.rows + 255) / 256), dim3(256), 0, s, A.rows,
                     A.ai.as<int>(), A.aj.as<int>(), v, d);
  LA_HIP(hipGetLastError());
  A.vals.set_from_device_f64(v, A.nnz, s);
}
}  // namespace laamd
