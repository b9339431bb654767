// Projection GEMM of the batched engine:
//   LS[n*ldc + m] = sum_k A[m*lda + k] * B[n*ldb + k]
// A = final eigenvector records (rows), B = projected half line strengths.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gain {

void gemm_projection(int precision, int M, int N, int K, const double* A,
                     int64_t lda, const double* B, int64_t ldb, double* C,
                     int64_t ldc, hipStream_t s);

// Zero-fill kernel (used instead of hipMemsetAsync so launch sequences stay
// capturable in HIP graphs).
void launch_fill_zero(double* p, int64_t n, hipStream_t s);

}  // namespace gain
