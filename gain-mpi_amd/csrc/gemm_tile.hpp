// One output tile of a matrix-core GEMM over FP64 operands, shared by the
// FP64 kernels (v_mfma_f64_16x16x4f64) and the FP32 ones
// (v_mfma_f32_16x16x4f32, FP64 accumulation).
//
// Global tiles are loaded as FP64 into registers one k step ahead and
// stored into LDS as the element type T (rounded to FP32 when T = float);
// operand fragments follow the instructions' lane layout: A(m,s) with
// m = lane%16, k slot s = lane/16; B(s,n) with n = lane%16, s = lane/16
// (slot s of product j of a k step holds k = 4j + s for FP64 and
// k = s BK/4 + j for FP32).  Accumulator element e of lane l holds
// D(m = l/16 + 4e, n = l%16) for the FP64 instruction and
// D(m = 4(l/16) + e, n = l%16) for the FP32 one.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "gemm.hpp"

namespace gain {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kPad = 1;

// Rows of output tiles walked together (see tile_order).
constexpr int kGroupRows = 8;

// FP32 sums are added into the FP64 accumulators every kF32Flush k.
constexpr int kF32Flush = 32;

// LDS layout of one stage, in elements of T: A as [BM][BK+PK] when A is
// k-contiguous in memory, else k-major [BK][BM+PT]; B as [BN][BK+PK].
// FP64: PK = 1, PT = 16 (the 4 k rows a fragment read touches fall in
// different bank groups).  FP32: PK = 4, so each lane reads its BK/4
// consecutive k of a row as 16-byte words (rows 16 bytes aligned; the 16
// rows of one read start in distinct 4-bank groups), and PT = 16/(BK/4), so
// the k-major rows one read touches, BK/4 apart, start 16 banks apart.  Two
// stages (one barrier per k step) when they fit in 80 KB, so two
// workgroups share a CU.
template <int BM, int BN, int BK, bool AK, typename T = double>
struct TileLds {
  static constexpr bool F32 = sizeof(T) == 4;
  static constexpr int PK = F32 ? 4 : kPad;
  static constexpr int kPadT = F32 ? 64 / BK : 16;
  static constexpr int SA = AK ? BM * (BK + PK) : BK * (BM + kPadT);
  static constexpr int SB = BN * (BK + PK);
  static constexpr bool DB = 2 * (SA + SB) * (int)sizeof(T) <= 80 * 1024;
  static constexpr int NB = DB ? 2 : 1;
  static constexpr int SIZE = NB * (SA + SB);
  static_assert(!F32 || (BK % 16 == 0 && SA % 4 == 0 && SB % 4 == 0),
                "FP32 stages need 16-byte rows");
  __device__ static int a_at(int m, int k) {
    return AK ? m * (BK + PK) + k : k * (BM + kPadT) + m;
  }
  __device__ static int b_at(int n, int k) { return n * (BK + PK) + k; }
};

// One BM x BN output tile, K range [kbeg, kend), WM x WN waves; each wave
// owns (BM/WM) x (BN/WN) = MT x NT MFMA tiles of 16 x 16.  T = double: FP64
// products and sums.  T = float: operands rounded to FP32, products and
// sums of kF32Flush consecutive k (counted from kbeg) in FP32, those sums
// added in FP64.
template <int BM, int BN, int BK, int WM, int WN, bool AK, typename T = double,
          int TH = WM * WN * 64>
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, const double* A,
                                          double* C, int m0, int n0,
                                          int kbeg, int kend, bool vecA,
                                          bool vecB, T* lds) {
  using L = TileLds<BM, BN, BK, AK, T>;
  constexpr bool F32 = std::is_same<T, float>::value;
  static_assert(F32 || std::is_same<T, double>::value, "element type");
  static_assert(!F32 || (BK <= kF32Flush && kF32Flush % BK == 0),
                "FP32 k step must divide the flush interval");
  static_assert(WM * WN * 64 == TH, "waves x 64 threads");
  constexpr int MT = BM / WM / 16, NT = BN / WN / 16;
  constexpr int QA = BM * BK / 4 / TH, QB = BN * BK / 4 / TH;
  static_assert(QA >= 1 && QB >= 1, "tile too small for the thread count");
  static_assert(QA * 4 * TH == BM * BK && QB * 4 * TH == BN * BK,
                "tile not a multiple of the thread count");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wmi = wave % WM, wni = wave / WM;
  d4 acc[MT][NT];
  f4 acc32[MT][NT];  // FP32 partial sums (T = float only)
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      acc32[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    }

  double ra[QA][4], rb[QB][4];
  auto load = [&](int k0) {
    const bool fullk = k0 + BK <= kend;
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int gi = tid + TH * q;
      if (AK) {  // 4 consecutive k of row m
        const int m = m0 + gi / (BK / 4), k = k0 + (gi % (BK / 4)) * 4;
        const double* p = A + (int64_t)m * g.lda + k;
        if (vecA && fullk && m < g.M) {
          const dbl2 v0 = *(const dbl2*)p, v1 = *(const dbl2*)(p + 2);
          ra[q][0] = v0.x; ra[q][1] = v0.y; ra[q][2] = v1.x; ra[q][3] = v1.y;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            ra[q][e] = (m < g.M && k + e < kend) ? p[e] : 0.0;
        }
      } else {  // 4 consecutive m of row k
        const int k = k0 + gi / (BM / 4), m = m0 + (gi % (BM / 4)) * 4;
        const double* p = A + (int64_t)k * g.lda + m;
        if (vecA && k < kend && m + 3 < g.M) {
          const dbl2 v0 = *(const dbl2*)p, v1 = *(const dbl2*)(p + 2);
          ra[q][0] = v0.x; ra[q][1] = v0.y; ra[q][2] = v1.x; ra[q][3] = v1.y;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            ra[q][e] = (k < kend && m + e < g.M) ? p[e] : 0.0;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int gi = tid + TH * q;
      const int n = n0 + gi / (BK / 4), k = k0 + (gi % (BK / 4)) * 4;
      const double* p = g.B + (int64_t)n * g.ldb + k;
      if (vecB && fullk && n < g.N) {
        const dbl2 v0 = *(const dbl2*)p, v1 = *(const dbl2*)(p + 2);
        rb[q][0] = v0.x; rb[q][1] = v0.y; rb[q][2] = v1.x; rb[q][3] = v1.y;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          rb[q][e] = (n < g.N && k + e < kend) ? p[e] : 0.0;
      }
    }
  };
  auto store = [&](int buf) {
    T* As = lds + buf * (L::SA + L::SB);
    T* Bs = As + L::SA;
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int gi = tid + TH * q;
      if constexpr (AK && F32) {  // 4 consecutive k: one 16-byte word
        *(f4*)(As + L::a_at(gi / (BK / 4), (gi % (BK / 4)) * 4)) =
            f4{(float)ra[q][0], (float)ra[q][1], (float)ra[q][2],
               (float)ra[q][3]};
        continue;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (AK)
          As[L::a_at(gi / (BK / 4), (gi % (BK / 4)) * 4 + e)] = (T)ra[q][e];
        else
          As[L::a_at((gi % (BM / 4)) * 4 + e, gi / (BM / 4))] = (T)ra[q][e];
      }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int gi = tid + TH * q;
      if constexpr (F32) {
        *(f4*)(Bs + L::b_at(gi / (BK / 4), (gi % (BK / 4)) * 4)) =
            f4{(float)rb[q][0], (float)rb[q][1], (float)rb[q][2],
               (float)rb[q][3]};
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          Bs[L::b_at(gi / (BK / 4), (gi % (BK / 4)) * 4 + e)] = rb[q][e];
      }
    }
  };
  auto compute = [&](int buf) {
    const T* As = lds + buf * (L::SA + L::SB);
    const T* Bs = As + L::SA;
    if constexpr (F32) {
      // The k of a step are dealt to the 4 lane groups in runs of BK/4:
      // group q = lane/16 feeds k = q BK/4 + j to product j of the step
      // (A and B alike, so each product still pairs equal k), and reads
      // them as 16-byte words.
      constexpr int KQ = BK / 4;
      const int kq0 = (lane >> 4) * KQ;
#pragma unroll
      for (int c = 0; c < KQ / 4; ++c) {
        f4 a[MT], b[NT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
          const int m = wmi * (BM / WM) + mt * 16 + (lane & 15);
          if (AK) {
            a[mt] = *(const f4*)(As + L::a_at(m, kq0 + 4 * c));
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              a[mt][e] = As[L::a_at(m, kq0 + 4 * c + e)];
          }
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          b[nt] = *(const f4*)(Bs + L::b_at(
                                        wni * (BN / WN) + nt * 16 + (lane & 15),
                                        kq0 + 4 * c));
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
              acc32[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                  a[mt][e], b[nt][e], acc32[mt][nt], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < BK / 4; ++kk) {
        const int kc = kk * 4 + (lane >> 4);
        T a[MT], b[NT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
          a[mt] = As[L::a_at(wmi * (BM / WM) + mt * 16 + (lane & 15), kc)];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
          b[nt] = Bs[L::b_at(wni * (BN / WN) + nt * 16 + (lane & 15), kc)];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
          for (int nt = 0; nt < NT; ++nt)
            acc[mt][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                a[mt], b[nt], acc[mt][nt], 0, 0, 0);
      }
    }
  };
  // T = float: after the k step ending at k0 + BK, add the FP32 sums into
  // the FP64 accumulators when kF32Flush k have been summed or K ends
  auto flush = [&](int k0) {
    if constexpr (F32) {
      if ((k0 + BK - kbeg) % kF32Flush == 0 || k0 + BK >= kend) {
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
          for (int j = 0; j < NT; ++j) {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              acc[i][j][e] += (double)acc32[i][j][e];
            acc32[i][j] = f4{0.f, 0.f, 0.f, 0.f};
          }
      }
    }
  };

  if constexpr (L::DB) {
    // stage s is read while stage s^1 is filled; the one barrier per step
    // orders this step's reads before the next step's overwrite
    if (kbeg < kend) {
      load(kbeg);
      store(0);
    }
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
      const bool more = k0 + BK < kend;
      if (more) load(k0 + BK);
      compute(buf);
      flush(k0);
      if (more) store(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  } else {
    if (kbeg < kend) load(kbeg);
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
      __syncthreads();
      store(0);
      __syncthreads();
      if (k0 + BK < kend) load(k0 + BK);
      compute(0);
      flush(k0);
    }
  }
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const int n = n0 + wni * (BN / WN) + nt * 16 + (lane & 15);
      if (n >= g.N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int mr = F32 ? 4 * (lane >> 4) + e : (lane >> 4) + 4 * e;
        const int m = m0 + wmi * (BM / WM) + mt * 16 + mr;
        if (m < g.M) C[(int64_t)n * g.ldc + m] = acc[mt][nt][e];
      }
    }
}

}  // namespace gain
