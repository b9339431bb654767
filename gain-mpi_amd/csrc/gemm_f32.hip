// FP32 matrix-core GEMM with FP64 accumulation (precision mode FP32).
//
//   C[z][n*ldc + m] = sum_k A(m,k) B(n,k)
//
// FP64 operands are rounded to FP32 when staged into LDS; products run on
// v_mfma_f32_16x16x4f32 and the FP32 sums of every 32 consecutive k are
// added into FP64 accumulators, so accumulation error stays at the FP64
// level and the error is the FP32 rounding of the operands (about 2^-24
// relative per element).  Each workgroup computes one output tile with
// gemm_tile (global tiles prefetched into registers one k step ahead,
// double-buffered LDS); tiles are dealt to the XCDs in contiguous runs of
// the grouped tile order, as in the FP64 GEMM.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "common.hpp"
#include "gemm.hpp"
#include "gemm_tile.hpp"

namespace gain {

template <int BM, int BN, int BK, int WM, int WN, bool AK>
__global__ __launch_bounds__(WM * WN * 64) void k_gemm_f32acc(GemmArgs g,
                                                              int ntn,
                                                              bool vecA,
                                                              bool vecB) {
  __shared__ __attribute__((aligned(16)))
  float lds[TileLds<BM, BN, BK, AK, float>::SIZE];
  const int nxy = ntn * gridDim.y;
  const int lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const int p = xcd_remap(lin, nxy * gridDim.z);
  const int z = p / nxy;
  int tm, tn;
  tile_order(p - z * nxy, gridDim.y, ntn, kGroupRows, &tm, &tn);
  gemm_tile<BM, BN, BK, WM, WN, AK, float>(
      g, g.A + (int64_t)z * g.sAz, g.C + (int64_t)z * g.sCz, tm * BM, tn * BN,
      0, g.K, vecA, vecB, lds);
}

struct F32Cfg {
  int bm, bn, bk, wm, wn;
};
// Tile shapes (index = configuration id of gemm_f32acc).
constexpr F32Cfg kF32Cfgs[kF32Configs] = {{128, 128, 32, 2, 4},
                                          {128, 128, 16, 2, 4},
                                          {64, 128, 32, 1, 4},
                                          {64, 128, 16, 1, 4}};
constexpr int kF32DefaultCfg = 0;

template <int BM, int BN, int BK, int WM, int WN>
static void launch_f32(const GemmArgs& g, hipStream_t s, bool vecA,
                       bool vecB) {
  const int ntm = (g.M + BM - 1) / BM, ntn = (g.N + BN - 1) / BN;
  GAIN_CHECK_ARG(ntm < 65536 && g.nz < 65536 &&
                     (int64_t)ntm * ntn * g.nz < (1ll << 31),
                 "gemm_f32: grid too large");
  const dim3 grid(ntn, ntm, g.nz);
  if (g.a_kcontig)
    hipLaunchKernelGGL((k_gemm_f32acc<BM, BN, BK, WM, WN, true>), grid,
                       dim3(WM * WN * 64), 0, s, g, ntn, vecA, vecB);
  else
    hipLaunchKernelGGL((k_gemm_f32acc<BM, BN, BK, WM, WN, false>), grid,
                       dim3(WM * WN * 64), 0, s, g, ntn, vecA, vecB);
}

void gemm_f32acc(const GemmArgs& g, hipStream_t s, int cfg) {
  if (g.M <= 0 || g.N <= 0) return;
  GAIN_CHECK_ARG(g.nz >= 1 && g.K >= 0, "gemm_f32: bad nz/K");
  if (cfg < 0) cfg = kF32DefaultCfg;
  GAIN_CHECK_ARG(cfg < kF32Configs, "gemm_f32: bad configuration %d", cfg);
  const bool vecA = (g.lda % 2 == 0) && ((uintptr_t)g.A % 16 == 0) &&
                    (g.sAz % 2 == 0);
  const bool vecB = (g.ldb % 2 == 0) && ((uintptr_t)g.B % 16 == 0);
  switch (cfg) {
    case 0: launch_f32<128, 128, 32, 2, 4>(g, s, vecA, vecB); break;
    case 1: launch_f32<128, 128, 16, 2, 4>(g, s, vecA, vecB); break;
    case 2: launch_f32<64, 128, 32, 1, 4>(g, s, vecA, vecB); break;
    case 3: launch_f32<64, 128, 16, 1, 4>(g, s, vecA, vecB); break;
  }
  check_launch("gemm_f32acc");
}

const char* gemm_f32_config_name(int cfg) {
  static char buf[kF32Configs][32];
  if (cfg < 0) cfg = kF32DefaultCfg;
  if (cfg >= kF32Configs) return "?";
  const F32Cfg& t = kF32Cfgs[cfg];
  snprintf(buf[cfg], sizeof(buf[cfg]), "%dx%dx%d/%dw", t.bm, t.bn, t.bk,
           t.wm * t.wn);
  return buf[cfg];
}

}  // namespace gain
