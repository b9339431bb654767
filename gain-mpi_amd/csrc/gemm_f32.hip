// FP32 matrix-core GEMM with FP64 accumulation (precision mode FP32).
//
//   C[z][n*ldc + m] = sum_k A(m,k) B(n,k)
//
// FP64 operands are rounded to FP32 when staged into LDS; products run on
// v_mfma_f32_16x16x4_f32 (A(m,k): m = lane%16, k = lane/16; B(k,n):
// n = lane%16; accumulator element e of lane l holds D(m = 4*(l/16) + e,
// n = l%16)).  The FP32 accumulators are added into FP64 ones every 32 k,
// so accumulation error stays at the FP64 level; the error is the FP32
// rounding of the operands (about 2^-24 relative per element).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "gemm.hpp"

namespace gain {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));

// Workgroup tile 64 x 128, K step 16, 4 waves of 64 x 32 (4 x 2 tiles).
template <bool AK>
__global__ __launch_bounds__(256) void k_gemm_f32acc(GemmArgs g, int ntn) {
  constexpr int BM = 64, BN = 128, BK = 16, MT = 4, NT = 2;
  __shared__ float As[BM][BK + 1];
  __shared__ float Bs[BN][BK + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tn = blockIdx.x % ntn, tm = blockIdx.x / ntn;
  const int m0 = tm * BM, n0 = tn * BN;
  const int z = blockIdx.z;
  const double* A = g.A + (int64_t)z * g.sAz;
  double* C = g.C + (int64_t)z * g.sCz;
  f4 acc[MT][NT];
  double acc64[MT][NT][4];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < 4; ++e) acc64[i][j][e] = 0.0;
    }
  int stage = 0;
  for (int k0 = 0; k0 < g.K; k0 += BK) {
    __syncthreads();
    // A: 64 x 16 (one group of 4 per thread); B: 128 x 16 (two groups)
    {
      if (AK) {
        const int m = m0 + (tid >> 2), k = k0 + (tid & 3) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          As[tid >> 2][(tid & 3) * 4 + e] =
              (m < g.M && k + e < g.K) ? (float)A[(int64_t)m * g.lda + k + e]
                                       : 0.f;
      } else {
        const int k = k0 + (tid >> 4), m = m0 + (tid & 15) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          As[(tid & 15) * 4 + e][tid >> 4] =
              (k < g.K && m + e < g.M) ? (float)A[(int64_t)k * g.lda + m + e]
                                       : 0.f;
      }
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int gi = tid + 256 * q;
        const int n = n0 + (gi >> 2), k = k0 + (gi & 3) * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          Bs[gi >> 2][(gi & 3) * 4 + e] =
              (n < g.N && k + e < g.K) ? (float)g.B[(int64_t)n * g.ldb + k + e]
                                       : 0.f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kc = kk * 4 + (lane >> 4);
      float a[MT], b[NT];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) a[mt] = As[mt * 16 + (lane & 15)][kc];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        b[nt] = Bs[wave * 32 + nt * 16 + (lane & 15)][kc];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(
              a[mt], b[nt], acc[mt][nt], 0, 0, 0);
    }
    if (++stage == 2 || k0 + BK >= g.K) {  // every 32 k and at the end
      stage = 0;
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) {
#pragma unroll
          for (int e = 0; e < 4; ++e) acc64[i][j][e] += (double)acc[i][j][e];
          acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
  }
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int n = n0 + wave * 32 + nt * 16 + (lane & 15);
      if (n >= g.N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = m0 + mt * 16 + 4 * (lane >> 4) + e;
        if (m < g.M) C[(int64_t)n * g.ldc + m] = acc64[mt][nt][e];
      }
    }
}

void gemm_f32acc(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  const int ntm = (g.M + 63) / 64, ntn = (g.N + 127) / 128;
  GAIN_CHECK_ARG((int64_t)ntm * ntn < (1ll << 31) && g.nz < 65536,
                 "gemm_f32: grid too large");
  dim3 grid(ntm * ntn, 1, g.nz);
  if (g.a_kcontig)
    hipLaunchKernelGGL((k_gemm_f32acc<true>), grid, dim3(256), 0, s, g, ntn);
  else
    hipLaunchKernelGGL((k_gemm_f32acc<false>), grid, dim3(256), 0, s, g, ntn);
  check_launch("gemm_f32acc");
}

}  // namespace gain
