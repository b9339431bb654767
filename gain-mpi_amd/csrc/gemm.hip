// FP64 GEMM on gfx950 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Tile shapes are template parameters chosen per GEMM shape by the engine's
// autotuner (kCfgs).  Operand
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

constexpr int kPad = 1;

// Rows of output tiles walked together (see tile_order).
constexpr int kGroupRows = 8;

// LDS layout of one stage: A as [BM][BK+1] when A is k-contiguous in
// memory, else k-major [BK][BM+16] (rows 16 doubles apart in bank space, so
// the 4 k rows a fragment read touches fall in different bank halves); B as
// [BN][BK+1].  Two stages (one barrier per k step) when they fit in 80 KB,
// so two workgroups share a CU.
template <int BM, int BN, int BK, bool AK>
struct TileLds {
  static constexpr int kPadT = 16;
  static constexpr int SA = AK ? BM * (BK + kPad) : BK * (BM + kPadT);
  static constexpr int SB = BN * (BK + kPad);
  static constexpr bool DB = 2 * (SA + SB) * 8 <= 80 * 1024;
  static constexpr int NB = DB ? 2 : 1;
  static constexpr int SIZE = NB * (SA + SB);
  __device__ static int a_at(int m, int k) {
    return AK ? m * (BK + kPad) + k : k * (BM + kPadT) + m;
  }
  __device__ static int b_at(int n, int k) { return n * (BK + kPad) + k; }
};

// One BM x BN output tile, K range [kbeg, kend), WM x WN waves; each wave
// owns (BM/WM) x (BN/WN) = MT x NT MFMA tiles of 16 x 16.  Global tiles are
// loaded into registers one k step ahead.
template <int BM, int BN, int BK, int WM, int WN, bool AK,
          int TH = WM * WN * 64>
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, const double* A,
                                          double* C, int m0, int n0,
                                          int kbeg, int kend, bool vecA,
                                          bool vecB, double* lds) {
  using L = TileLds<BM, BN, BK, AK>;
  static_assert(WM * WN * 64 == TH, "waves x 64 threads");
  constexpr int MT = BM / WM / 16, NT = BN / WN / 16;
  constexpr int QA = BM * BK / 4 / TH, QB = BN * BK / 4 / TH;
  static_assert(QA >= 1 && QB >= 1, "tile too small for the thread count");
  static_assert(QA * 4 * TH == BM * BK && QB * 4 * TH == BN * BK,
                "tile not a multiple of the thread count");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wmi = wave % WM, wni = wave / WM;
  d4 acc[MT][NT];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

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
    double* As = lds + buf * (L::SA + L::SB);
    double* Bs = As + L::SA;
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int gi = tid + TH * q;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (AK)
          As[L::a_at(gi / (BK / 4), (gi % (BK / 4)) * 4 + e)] = ra[q][e];
        else
          As[L::a_at((gi % (BM / 4)) * 4 + e, gi / (BM / 4))] = ra[q][e];
      }
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int gi = tid + TH * q;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        Bs[L::b_at(gi / (BK / 4), (gi % (BK / 4)) * 4 + e)] = rb[q][e];
    }
  };
  auto compute = [&](int buf) {
    const double* As = lds + buf * (L::SA + L::SB);
    const double* Bs = As + L::SA;
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kc = kk * 4 + (lane >> 4);
      double a[MT], b[NT];
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
  };

  if (L::DB) {
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
        const int m = m0 + wmi * (BM / WM) + mt * 16 + (lane >> 4) + 4 * e;
        if (m < g.M) C[(int64_t)n * g.ldc + m] = acc[mt][nt][e];
      }
    }
}

template <int BM, int BN, int BK, int WM, int WN, bool AK>
__global__ __launch_bounds__(WM * WN * 64) void k_gemm_f64(GemmArgs g,
                                                           int kchunk,
                                                  int ntn, bool vecA,
                                                  bool vecB) {
  __shared__ double lds[TileLds<BM, BN, BK, AK>::SIZE];
  // linear id -> (z slice, split, tile): each XCD takes a contiguous run
  // of the grouped tile order, so the workgroups resident on one XCD cover
  // a compact block of rows and columns and share operand slabs in its L2
  const int nxy = ntn * gridDim.y;
  const int lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int p = xcd_remap(lin, nxy * gridDim.z);
  const int zk = p / nxy;
  int tm, tn;
  tile_order(p - zk * nxy, gridDim.y, ntn, kGroupRows, &tm, &tn);
  const int z = zk % g.nz, ks = zk / g.nz;
  const double* A = g.A + (int64_t)z * g.sAz;
  double* C = g.C + (int64_t)z * g.sCz + (int64_t)ks * g.sCks;
  const int kbeg = ks * kchunk;
  const int kend = min(g.K, kbeg + kchunk);
  gemm_tile<BM, BN, BK, WM, WN, AK>(g, A, C, tm * BM, tn * BN, kbeg, kend,
                                    vecA, vecB, lds);
}

// Persistent launch over the tiles of several GEMM problems (one workgroup
// walks the tile list with a fixed stride, so the assignment is static and
// the launch needs no counters).
template <int BM, int BN, int BK, int WM, int WN, bool AK>
__global__ __launch_bounds__(WM * WN * 64) void k_gemm_f64_grouped(
    const GemmArgs* __restrict__ probs, const GroupTile* __restrict__ tiles,
    int ntiles) {
  __shared__ double lds[TileLds<BM, BN, BK, AK>::SIZE];
  for (int it = blockIdx.x; it < ntiles; it += gridDim.x) {
    const GroupTile tl = tiles[it];
    const GemmArgs& g = probs[tl.prob];
    const double* A = g.A + (int64_t)tl.z * g.sAz;
    double* C = g.C + (int64_t)tl.z * g.sCz;
    const bool vecA = (g.lda % 2 == 0) && ((uintptr_t)A % 16 == 0);
    const bool vecB = (g.ldb % 2 == 0) && ((uintptr_t)g.B % 16 == 0);
    __syncthreads();  // LDS of the previous tile fully consumed
    gemm_tile<BM, BN, BK, WM, WN, AK>(g, A, C, tl.tm * BM, tl.tn * BN, 0,
                                      g.K, vecA, vecB, lds);
  }
}

struct TileCfg {
  int bm, bn, bk, waves;
};
// The tile shapes the autotuner chooses from (index = configuration id).
constexpr TileCfg kCfgs[kGemmConfigs] = {
    {64, 64, 16, 4},    {64, 128, 16, 4},   {128, 128, 16, 4},
    {64, 256, 16, 4},   {128, 64, 16, 4},   {64, 128, 32, 4},
    {128, 128, 32, 4},  {128, 256, 16, 4},  {256, 128, 16, 4},
    {128, 256, 32, 4},  {128, 256, 16, 8},  {256, 128, 16, 8},
    {128, 128, 16, 8},  {128, 256, 32, 8},  {64, 256, 32, 8}};

template <int BM, int BN, int BK, int WM, int WN>
static void launch_cfg(const GemmArgs& g, hipStream_t s, int kchunk,
                       bool vecA, bool vecB) {
  const int ntm = (g.M + BM - 1) / BM, ntn = (g.N + BN - 1) / BN;
  dim3 grid(ntn, ntm, g.nz * g.ksplit);
  GAIN_CHECK_ARG((int64_t)grid.z < 65536 && (int64_t)ntm < 65536 &&
                     (int64_t)ntn * ntm * grid.z < (1ll << 31),
                 "gemm: grid too large");
  if (g.a_kcontig)
    hipLaunchKernelGGL((k_gemm_f64<BM, BN, BK, WM, WN, true>), grid,
                       dim3(WM * WN * 64), 0, s, g, kchunk, ntn, vecA, vecB);
  else
    hipLaunchKernelGGL((k_gemm_f64<BM, BN, BK, WM, WN, false>), grid,
                       dim3(WM * WN * 64), 0, s, g, kchunk, ntn, vecA, vecB);
}

int gemm_default_config(int N) { return N > 64 ? 1 : 0; }

void gemm_f64(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  GAIN_CHECK_ARG(g.nz >= 1 && g.ksplit >= 1, "gemm: bad nz/ksplit");
  const int cfg = g.cfg < 0 ? gemm_default_config(g.N) : g.cfg;
  GAIN_CHECK_ARG(cfg < kGemmConfigs, "gemm: bad configuration %d", cfg);
  const int bk = kCfgs[cfg].bk;
  int kchunk = (g.K + g.ksplit - 1) / g.ksplit;
  kchunk = (kchunk + bk - 1) / bk * bk;
  const bool vecA = (g.lda % 2 == 0) && ((uintptr_t)g.A % 16 == 0) &&
                    (g.sAz % 2 == 0);
  const bool vecB = (g.ldb % 2 == 0) && ((uintptr_t)g.B % 16 == 0);
  switch (cfg) {
    case 0: launch_cfg<64, 64, 16, 2, 2>(g, s, kchunk, vecA, vecB); break;
    case 1: launch_cfg<64, 128, 16, 1, 4>(g, s, kchunk, vecA, vecB); break;
    case 2: launch_cfg<128, 128, 16, 2, 2>(g, s, kchunk, vecA, vecB); break;
    case 3: launch_cfg<64, 256, 16, 1, 4>(g, s, kchunk, vecA, vecB); break;
    case 4: launch_cfg<128, 64, 16, 2, 2>(g, s, kchunk, vecA, vecB); break;
    case 5: launch_cfg<64, 128, 32, 1, 4>(g, s, kchunk, vecA, vecB); break;
    case 6: launch_cfg<128, 128, 32, 2, 2>(g, s, kchunk, vecA, vecB); break;
    case 7: launch_cfg<128, 256, 16, 2, 2>(g, s, kchunk, vecA, vecB); break;
    case 8: launch_cfg<256, 128, 16, 2, 2>(g, s, kchunk, vecA, vecB); break;
    case 9: launch_cfg<128, 256, 32, 2, 2>(g, s, kchunk, vecA, vecB); break;
    case 10: launch_cfg<128, 256, 16, 2, 4>(g, s, kchunk, vecA, vecB); break;
    case 11: launch_cfg<256, 128, 16, 4, 2>(g, s, kchunk, vecA, vecB); break;
    case 12: launch_cfg<128, 128, 16, 2, 4>(g, s, kchunk, vecA, vecB); break;
    case 13: launch_cfg<128, 256, 32, 2, 4>(g, s, kchunk, vecA, vecB); break;
    case 14: launch_cfg<64, 256, 32, 1, 8>(g, s, kchunk, vecA, vecB); break;
  }
  check_launch("gemm_f64");
}

template <int BM, int BN, int BK, int WM, int WN>
static void launch_grouped_cfg(const GemmArgs* probs, const GroupTile* tiles,
                               int ntiles, bool ak, int grid, hipStream_t s) {
  if (ak)
    hipLaunchKernelGGL((k_gemm_f64_grouped<BM, BN, BK, WM, WN, true>),
                       dim3(grid), dim3(WM * WN * 64), 0, s, probs, tiles,
                       ntiles);
  else
    hipLaunchKernelGGL((k_gemm_f64_grouped<BM, BN, BK, WM, WN, false>),
                       dim3(grid), dim3(WM * WN * 64), 0, s, probs, tiles,
                       ntiles);
}

void gemm_f64_grouped(const GemmArgs* probs, const GroupTile* tiles,
                      int ntiles, bool a_kcontig, int cfg, int grid,
                      hipStream_t s) {
  if (ntiles <= 0) return;
  GAIN_CHECK_ARG(cfg >= 0 && cfg < kGemmConfigs, "gemm: bad configuration");
  grid = std::max(1, std::min(grid, ntiles));
  switch (cfg) {
    case 0: launch_grouped_cfg<64, 64, 16, 2, 2>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 1: launch_grouped_cfg<64, 128, 16, 1, 4>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 2: launch_grouped_cfg<128, 128, 16, 2, 2>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 3: launch_grouped_cfg<64, 256, 16, 1, 4>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 4: launch_grouped_cfg<128, 64, 16, 2, 2>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 5: launch_grouped_cfg<64, 128, 32, 1, 4>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 6: launch_grouped_cfg<128, 128, 32, 2, 2>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 7: launch_grouped_cfg<128, 256, 16, 2, 2>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 8: launch_grouped_cfg<256, 128, 16, 2, 2>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 9: launch_grouped_cfg<128, 256, 32, 2, 2>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 10: launch_grouped_cfg<128, 256, 16, 2, 4>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 11: launch_grouped_cfg<256, 128, 16, 4, 2>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 12: launch_grouped_cfg<128, 128, 16, 2, 4>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 13: launch_grouped_cfg<128, 256, 32, 2, 4>(probs, tiles, ntiles, a_kcontig, grid, s); break;
    case 14: launch_grouped_cfg<64, 256, 32, 1, 8>(probs, tiles, ntiles, a_kcontig, grid, s); break;
  }
  check_launch("gemm_f64_grouped");
}

int64_t gemm_grid_items(const GemmArgs& g) {
  const int cfg = g.cfg < 0 ? gemm_default_config(g.N) : g.cfg;
  const TileCfg& t = kCfgs[cfg];
  return (int64_t)((g.M + t.bm - 1) / t.bm) * ((g.N + t.bn - 1) / t.bn) *
         g.nz * g.ksplit * t.waves * 64;
}

int gemm_tile_m(int cfg) { return kCfgs[cfg].bm; }
int gemm_tile_n(int cfg) { return kCfgs[cfg].bn; }

const char* gemm_config_name(int cfg) {
  static char buf[kGemmConfigs][32];
  if (cfg < 0 || cfg >= kGemmConfigs) return "default";
  snprintf(buf[cfg], sizeof(buf[cfg]), "%dx%dx%d/%dw", kCfgs[cfg].bm,
           kCfgs[cfg].bn, kCfgs[cfg].bk, kCfgs[cfg].waves);
  return buf[cfg];
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
    case GAIN_PREC_FP32: {
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
      gemm_f32acc(g, s);
      return;
    }
    default:
      fail(GAIN_ERR_ARG, "precision %d not available in this build",
           precision);
  }
}

}  // namespace gain
