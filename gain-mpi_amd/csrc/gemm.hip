// Projection GEMM dispatch for the batched engine.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace gain {

__global__ __launch_bounds__(256) void k_fill_zero(double* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

void launch_fill_zero(double* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)),
                     dim3(256), 0, s, p, n);
  check_launch("fill_zero");
}

void gemm_projection(int precision, int M, int N, int K, const double* A,
                     int64_t lda, const double* B, int64_t ldb, double* C,
                     int64_t ldc, hipStream_t s) {
  switch (precision) {
    case GAIN_PREC_FP64:
      launch_gemm_tn_valu(M, N, K, A, lda, B, ldb, C, ldc, s);
      return;
    default:
      fail(GAIN_ERR_ARG, "precision %d not available in this build",
           precision);
  }
}

}  // namespace gain
