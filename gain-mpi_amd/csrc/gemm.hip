// FP64 GEMM on gfx950 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Tile shapes are template parameters chosen per GEMM shape by the engine's
// autotuner (kCfgs); each workgroup computes one output tile with
// gemm_tile (global tiles prefetched into registers while the previous one
// is consumed from LDS).  Workgroups that share an A tile are placed on the
// same XCD (its own L2).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "gemm.hpp"
#include "gemm_tile.hpp"
#include "kernels.hpp"

namespace gain {

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
                     int64_t ldc, hipStream_t s, int f32_cfg) {
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
      gemm_f32acc(g, s, f32_cfg);
      return;
    }
    default:
      fail(GAIN_ERR_ARG, "precision %d not available in this build",
           precision);
  }
}

}  // namespace gain
