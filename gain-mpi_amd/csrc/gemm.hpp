// Dense GEMMs of the engine (gfx950 MFMA).
//
//   C[z][ks][n*ldc + m] = sum_{k in split ks} A[z](m,k) * B(n,k)
//
// A is either K-contiguous (A[m*lda + k], eigenvector records) or
// M-contiguous (A[k*lda + m], the dipole mu_alpha with v_F fastest);
// B is always K-contiguous (B[n*ldb + k]); C is M-contiguous.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gain {

struct GemmArgs {
  int M, N, K;
  const double* A;
  int64_t lda;
  int64_t sAz;    // A stride between z slices
  bool a_kcontig;
  const double* B;
  int64_t ldb;
  double* C;
  int64_t ldc;
  int64_t sCz;    // C stride between z slices
  int64_t sCks;   // C stride between split-K partials
  int nz;         // number of z slices (e.g. the 3 dipole components)
  int ksplit;     // split-K partial count (partials are summed by the caller)
  int cfg = -1;   // tile configuration (-1: default for the shape)
};

constexpr int kGemmConfigs = 15;

// Bijective remap of a linear workgroup id in [0, total): the hardware deals
// ids round-robin over the 8 XCDs (XCD = lin % 8); XCD x receives the
// contiguous position range starting at x*(total/8) + min(x, total%8).
__host__ __device__ inline int xcd_remap(int lin, int total) {
  const int q = total / 8, r = total % 8, x = lin % 8;
  return x * q + (x < r ? x : r) + lin / 8;
}

// Position p in [0, ntm*ntn) -> output tile (tm, tn): groups of `gm` tile
// rows, walked column by column inside a group, so consecutive positions
// share A rows (within a column) and B columns (across the group).
__host__ __device__ inline void tile_order(int p, int ntm, int ntn, int gm,
                                           int* tm, int* tn) {
  const int per = gm * ntn;
  const int grp = p / per;
  const int first = grp * gm;
  const int rows = ntm - first < gm ? ntm - first : gm;
  const int r = p - grp * per;
  *tm = first + r % rows;
  *tn = r / rows;
}

// One output tile of one problem in a grouped (persistent) launch.
struct GroupTile {
  int prob, tm, tn, z;
};
// Tiles of several problems whose A operands share one layout, computed by
// `grid` persistent workgroups; probs and tiles are device arrays.  Split-K
// is not used (ksplit ignored).
void gemm_f64_grouped(const GemmArgs* probs, const GroupTile* tiles,
                      int ntiles, bool a_kcontig, int cfg, int grid,
                      hipStream_t s);
// FP32 matrix-core products with FP64 accumulation over FP64 operands
// (ksplit and g.cfg ignored): tile configuration cfg in [0, kF32Configs),
// -1 = default.
constexpr int kF32Configs = 4;
void gemm_f32acc(const GemmArgs& g, hipStream_t s, int cfg = -1);
const char* gemm_f32_config_name(int cfg);
// Work-items of the launch gemm_f64 makes for g (as profilers report it).
int64_t gemm_grid_items(const GemmArgs& g);
int gemm_tile_m(int cfg);
int gemm_tile_n(int cfg);
void gemm_f64(const GemmArgs& g, hipStream_t s);
int gemm_default_config(int N);
const char* gemm_config_name(int cfg);

// Projection GEMM of the batched engine at the requested precision (FP64
// or FP32; f32_cfg = the FP32 GEMM's tile configuration):
//   LS[n*ldc + m] = sum_k A[m*lda + k] * B[n*ldb + k]
void gemm_projection(int precision, int M, int N, int K, const double* A,
                     int64_t lda, const double* B, int64_t ldb, double* C,
                     int64_t ldc, hipStream_t s, int f32_cfg = -1);

// Zero-fill kernel (used instead of hipMemsetAsync so launch sequences stay
// capturable in HIP graphs).
void launch_fill_zero(double* p, int64_t n, hipStream_t s);

}  // namespace gain
