// FP64 GEMM on gfx950 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Workgroup tile 64 (M) x BN (N), K step 16, 4 waves splitting N.  Operand
// fragments follow the instruction's lane layout: A(m,k) with m = lane%16,
// k = lane/16; B(k,n) with n = lane%16, k = lane/16; accumulator element e
// of lane l holds D(m = l/16 + 4e, n = l%16).  Global tiles are prefetched
// into registers while the previous tile is consumed from LDS.  Workgroups
// that share an A tile are placed on the same XCD (its own L2).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

namespace gain {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));

constexpr int kBM = 64, kBK = 16, kPad = 1;

// Bijective remap of the linear workgroup id so that each of the 8 XCDs
// (dispatch is round-robin over XCDs) receives a contiguous range of tiles.
__device__ inline int xcd_remap(int lin, int total) {
  if (total % 8) return lin;
  return (lin % 8) * (total / 8) + lin / 8;
}

template <int BN, bool AK>
__global__ __launch_bounds__(256) void k_gemm_f64(GemmArgs g, int kchunk,
                                                  int ntn, bool vecA,
                                                  bool vecB) {
  constexpr int NT = BN / 64;  // 16-wide column tiles per wave
  __shared__ double As[kBM][kBK + kPad];
  __shared__ double Bs[BN][kBK + kPad];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // tiles: n fastest so neighbours share the A tile
  const int nxy = ntn * gridDim.y;
  const int t = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, nxy);
  const int tn = t % ntn, tm = t / ntn;
  const int m0 = tm * kBM, n0 = tn * BN;
  const int z = blockIdx.z % g.nz, ks = blockIdx.z / g.nz;
  const double* A = g.A + (int64_t)z * g.sAz;
  double* C = g.C + (int64_t)z * g.sCz + (int64_t)ks * g.sCks;
  const int kbeg = ks * kchunk;
  const int kend = min(g.K, kbeg + kchunk);

  d4 acc[4][NT];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  double ra[4], rb[4 * NT];
  // A load coordinates
  const int ar = AK ? (tid >> 2) : ((tid & 15) * 4);  // m (AK) / m base (!AK)
  const int ak = AK ? ((tid & 3) * 4) : (tid >> 4);   // k base / k
  const int br = tid >> 2, bk = (tid & 3) * 4;

  auto load = [&](int k0) {
    const bool fullk = k0 + kBK <= kend;
    if (AK) {
      const int m = m0 + ar;
      const double* p = A + (int64_t)m * g.lda + k0 + ak;
      if (vecA && fullk && m < g.M) {
        const dbl2 v0 = *(const dbl2*)p, v1 = *(const dbl2*)(p + 2);
        ra[0] = v0.x; ra[1] = v0.y; ra[2] = v1.x; ra[3] = v1.y;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          ra[q] = (m < g.M && k0 + ak + q < kend) ? p[q] : 0.0;
      }
    } else {
      const int k = k0 + ak;
      const double* p = A + (int64_t)k * g.lda + m0 + ar;
      if (vecA && k < kend && m0 + ar + 3 < g.M) {
        const dbl2 v0 = *(const dbl2*)p, v1 = *(const dbl2*)(p + 2);
        ra[0] = v0.x; ra[1] = v0.y; ra[2] = v1.x; ra[3] = v1.y;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          ra[q] = (k < kend && m0 + ar + q < g.M) ? p[q] : 0.0;
      }
    }
#pragma unroll
    for (int pp = 0; pp < NT; ++pp) {
      const int n = n0 + pp * 64 + br;
      const double* p = g.B + (int64_t)n * g.ldb + k0 + bk;
      if (vecB && fullk && n < g.N) {
        const dbl2 v0 = *(const dbl2*)p, v1 = *(const dbl2*)(p + 2);
        rb[pp * 4 + 0] = v0.x; rb[pp * 4 + 1] = v0.y;
        rb[pp * 4 + 2] = v1.x; rb[pp * 4 + 3] = v1.y;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          rb[pp * 4 + q] = (n < g.N && k0 + bk + q < kend) ? p[q] : 0.0;
      }
    }
  };

  if (kbeg < kend) load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += kBK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (AK)
        As[ar][ak + q] = ra[q];
      else
        As[ar + q][ak] = ra[q];
    }
#pragma unroll
    for (int pp = 0; pp < NT; ++pp)
#pragma unroll
      for (int q = 0; q < 4; ++q) Bs[pp * 64 + br][bk + q] = rb[pp * 4 + q];
    __syncthreads();
    if (k0 + kBK < kend) load(k0 + kBK);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int kc = kk * 4 + (lane >> 4);
      double a[4], b[NT];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) a[mt] = As[mt * 16 + (lane & 15)][kc];
#pragma unroll
      for (int nt = 0; nt < NT; ++nt)
        b[nt] = Bs[wave * (BN / 4) + nt * 16 + (lane & 15)][kc];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              a[mt], b[nt], acc[mt][nt], 0, 0, 0);
    }
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int n = n0 + wave * (BN / 4) + nt * 16 + (lane & 15);
      if (n >= g.N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = m0 + mt * 16 + (lane >> 4) + 4 * e;
        if (m < g.M) C[(int64_t)n * g.ldc + m] = acc[mt][nt][e];
      }
    }
}

void gemm_f64(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  GAIN_CHECK_ARG(g.nz >= 1 && g.ksplit >= 1, "gemm: bad nz/ksplit");
  int kchunk = (g.K + g.ksplit - 1) / g.ksplit;
  kchunk = (kchunk + kBK - 1) / kBK * kBK;
  const bool vecA = (g.lda % 2 == 0) && ((uintptr_t)g.A % 16 == 0) &&
                    (g.sAz % 2 == 0);
  const bool vecB = (g.ldb % 2 == 0) && ((uintptr_t)g.B % 16 == 0);
  const int ntm = (g.M + kBM - 1) / kBM;
  const bool wide = g.N > 64;
  const int BN = wide ? 128 : 64;
  const int ntn = (g.N + BN - 1) / BN;
  dim3 grid(ntn, ntm, g.nz * g.ksplit);
  GAIN_CHECK_ARG((int64_t)grid.z < 65536 && (int64_t)ntm < 65536,
                 "gemm: grid too large");
  if (wide) {
    if (g.a_kcontig)
      hipLaunchKernelGGL((k_gemm_f64<128, true>), grid, dim3(256), 0, s, g,
                         kchunk, ntn, vecA, vecB);
    else
      hipLaunchKernelGGL((k_gemm_f64<128, false>), grid, dim3(256), 0, s, g,
                         kchunk, ntn, vecA, vecB);
  } else {
    if (g.a_kcontig)
      hipLaunchKernelGGL((k_gemm_f64<64, true>), grid, dim3(256), 0, s, g,
                         kchunk, ntn, vecA, vecB);
    else
      hipLaunchKernelGGL((k_gemm_f64<64, false>), grid, dim3(256), 0, s, g,
                         kchunk, ntn, vecA, vecB);
  }
  check_launch("gemm_f64");
}

// ---------------------------------------------------------------------------
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
    case GAIN_PREC_FP64: {
      GemmArgs g{};
      g.M = M;
      g.N = N;
      g.K = K;
      g.A = A;
      g.lda = lda;
      g.a_kcontig = true;
      g.B = B;
      g.ldb = ldb;
      g.C = C;
      g.ldc = ldc;
      g.nz = 1;
      g.ksplit = 1;
      gemm_f64(g, s);
      return;
    }
    default:
      fail(GAIN_ERR_ARG, "precision %d not available in this build",
           precision);
  }
}

}  // namespace gain
