// Split-INT8 slicing and sliced GEMM (v_mfma_i32_32x32x32_i8).  The scheme,
// its exactness bound and the tiled plane layout are described in
// ozaki.hpp.
//
// GEMM design (gfx950): 4 waves per workgroup, one wave per SIMD (the S
// level accumulators need more than 256 registers).  Each wave owns TM x TN
// output tiles of 32 x 32 and keeps S int32 accumulator sets for them.
// Operand fragments are read straight from the tiled planes into registers
// (one coalesced 1 KB block per fragment; the waves of a workgroup that
// share rows or columns hit the same lines in L1/L2), double-buffered one
// k block ahead, so the kernel has no LDS and no barriers.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "gemm.hpp"
#include "ozaki.hpp"

namespace gain {

typedef int i4v __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// One workgroup per padded row: row maximum, exponent, S slices written
// 4 bytes per thread into the tiled planes (zeros beyond the row / K).
template <int S>
__global__ __launch_bounds__(256) void k_slice_i8(
    const double* __restrict__ src, int64_t ld, int rows, int K, int64_t Kp,
    int64_t plane, int8_t* __restrict__ planes, int* __restrict__ expo) {
  __shared__ double red[4];
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  const bool live = r < rows;
  const double* x = src + (live ? (int64_t)r * ld : 0);
  double mx = 0.0;
  if (live)
    for (int k = tid; k < K; k += 256) mx = fmax(mx, fabs(x[k]));
  for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);  // mx = f 2^e, f in [0.5, 1)
  if (tid == 0) expo[r] = e;
  int8_t* rowbase = planes + (int64_t)(r >> 5) * (Kp << 5) + (r & 31) * 32;
  for (int64_t k0 = 4 * tid; k0 < Kp; k0 += 1024) {
    uint32_t w[S];
#pragma unroll
    for (int i = 0; i < S; ++i) w[i] = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t k = k0 + q;
      double v = (live && k < K) ? ldexp(x[k], -e) : 0.0;  // |v| < 1
#pragma unroll
      for (int i = 0; i < S; ++i) {
        // |v| <= 2^{-7i} here, so |n| <= 64; the subtraction is exact
        const double n = rint(ldexp(v, 6 + 7 * i));
        v -= ldexp(n, -6 - 7 * i);
        w[i] |= (uint32_t)(uint8_t)(int8_t)(int)n << (8 * q);
      }
    }
    int8_t* dst = rowbase + (k0 >> 5) * 1024 + (k0 & 31);
#pragma unroll
    for (int i = 0; i < S; ++i) *(uint32_t*)(dst + i * plane) = w[i];
  }
}

void slice_rows_i8(int S, const double* src, int64_t ld, int rows, int K,
                   int8_t* planes, int64_t plane, int* expo, hipStream_t s) {
  const int64_t Rp = i8_rows_pad(rows), Kp = i8_kpad(K);
  GAIN_CHECK_ARG(rows >= 0 && K > 0 && plane >= Rp * Kp,
                 "slice_rows_i8: plane of %lld bytes < %lld x %lld",
                 (long long)plane, (long long)Rp, (long long)Kp);
  GAIN_CHECK_ARG(((uintptr_t)planes & 3) == 0 && (plane & 3) == 0,
                 "slice_rows_i8: planes must be 4-byte aligned");
  switch (S) {
#define GAIN_SLICE_CASE(SS)                                                  \
  case SS:                                                                   \
    hipLaunchKernelGGL(k_slice_i8<SS>, dim3((unsigned)Rp), dim3(256), 0, s, \
                       src, ld, rows, K, Kp, plane, planes, expo);           \
    break;
    GAIN_SLICE_CASE(5)
    GAIN_SLICE_CASE(6)
    GAIN_SLICE_CASE(7)
#undef GAIN_SLICE_CASE
    default:
      fail(GAIN_ERR_ARG, "int8 slice count %d not supported (5..7)", S);
  }
  check_launch("slice_rows_i8");
}

// ---------------------------------------------------------------------------
// Tile configurations: wave tile (32 TM) x (32 TN), WM x WN waves.
struct I8Cfg {
  int tm, tn, wm, wn;
};
constexpr I8Cfg kI8Cfg[kI8Configs] = {
    {1, 2, 2, 2},  // 64 x 128
    {2, 1, 2, 2},  // 128 x 64
    {1, 2, 4, 1},  // 128 x 64
    {1, 2, 1, 4},  // 32 x 256
    {2, 1, 4, 1},  // 256 x 32
    {2, 1, 1, 4},  // 64 x 128
};

const char* i8_config_name(int cfg) {
  static const char* names[kI8Configs] = {
      "64x128/w32x64", "128x64/w64x32", "128x64/w32x64",
      "32x256/w32x64", "256x32/w64x32", "64x128/w64x32"};
  return cfg >= 0 && cfg < kI8Configs ? names[cfg] : "?";
}

template <int S, int TM, int TN, int WM, int WN>
__global__ __launch_bounds__(256) void k_gemm_i8(I8GemmArgs g, int ntm,
                                                 int ntn, int rb0,
                                                 int kb_per) {
  constexpr int BMB = TM * WM, BNB = TN * WN;  // tile size in 32-row blocks
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int ntiles = ntm * ntn;
  const int per_z = ntiles * g.ksplit;
  const int p = xcd_remap(blockIdx.x, gridDim.x);
  const int z = p / per_z;
  const int rem = p - z * per_z;
  const int ksi = rem / ntiles;
  int tm, tn;
  tile_order(rem - ksi * ntiles, ntm, ntn, 8, &tm, &tn);
  const int64_t nkb = g.Kp >> 5;
  const int kb0 = ksi * kb_per;
  const int kb1 = (int)min((int64_t)kb0 + kb_per, nkb);
  const int arb = rb0 + tm * BMB + wm * TM;  // A row block of m-tile 0
  const int brb = tn * BNB + wn * TN;        // B row block of n-tile 0
  const int lane_off = (lane & 31) * 32 + 16 * (lane >> 5);
  const int8_t* Ab = g.A + z * g.a_z + (int64_t)arb * nkb * 1024 + lane_off;
  const int8_t* Bb = g.B + (int64_t)brb * nkb * 1024 + lane_off;

  i16v acc[S][TM][TN];
#pragma unroll
  for (int L = 0; L < S; ++L)
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[L][a][b][e] = 0;

  i4v fa0[S][TM], fb0[S][TN], fa1[S][TM], fb1[S][TN];
  auto load = [&](i4v(&fa)[S][TM], i4v(&fb)[S][TN], int kb) {
#pragma unroll
    for (int i = 0; i < S; ++i) {
#pragma unroll
      for (int a = 0; a < TM; ++a)
        fa[i][a] = *(const i4v*)(Ab + i * g.a_plane +
                                 ((int64_t)a * nkb + kb) * 1024);
#pragma unroll
      for (int b = 0; b < TN; ++b)
        fb[i][b] = *(const i4v*)(Bb + i * g.b_plane +
                                 ((int64_t)b * nkb + kb) * 1024);
    }
  };
  auto mma = [&](const i4v(&fa)[S][TM], const i4v(&fb)[S][TN]) {
    // slice pairs (i, j) with i + j < S; consecutive products go to
    // different accumulators
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
      for (int j = 0; j < S - i; ++j)
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < TN; ++b)
            acc[i + j][a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                fa[i][a], fb[j][b], acc[i + j][a][b], 0, 0, 0);
  };

  int kb = kb0;
  if (kb < kb1) load(fa0, fb0, kb);
  for (; kb + 1 < kb1; kb += 2) {
    load(fa1, fb1, kb + 1);
    mma(fa0, fb0);
    if (kb + 2 < kb1) load(fa0, fb0, kb + 2);
    mma(fa1, fb1);
  }
  if (kb < kb1) mma(fa0, fb0);

  // epilogue: lane holds column n = lane%32 of each n-tile and rows
  // 8(e/4) + 4(lane/32) + e%4 of each m-tile
  double* C = g.C + z * g.c_z + ksi * g.c_ks;
  const int* eA = g.eA + z * g.ea_z;
#pragma unroll
  for (int b = 0; b < TN; ++b) {
    const int n = (brb + b) * 32 + (lane & 31);
    if (n >= g.N) continue;
    const int eb = g.eB[n] - 12;
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int ar = (arb + a) * 32 + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
        const int m = ar - g.row0;
        if (m < 0 || m >= g.M) continue;
        double v = (double)acc[S - 1][a][b][e];
#pragma unroll
        for (int L = S - 2; L >= 0; --L)
          v = (double)acc[L][a][b][e] + v * 0.0078125;  // 2^-7
        C[(int64_t)n * g.ldc + m] = ldexp(v, eA[ar] + eb);
      }
  }
}

namespace {

struct I8Launch {
  int ntm, ntn, rb0, kb_per;
  int64_t grid;
};

I8Launch i8_launch(const I8GemmArgs& g) {
  const I8Cfg& cf = kI8Cfg[g.cfg];
  const int bm = 32 * cf.tm * cf.wm, bn = 32 * cf.tn * cf.wn;
  I8Launch L;
  L.rb0 = g.row0 / 32;
  const int rows = g.row0 + g.M - L.rb0 * 32;
  L.ntm = (rows + bm - 1) / bm;
  L.ntn = (g.N + bn - 1) / bn;
  const int64_t nkb = g.Kp / 32;
  L.kb_per = (int)((nkb + g.ksplit - 1) / g.ksplit);
  L.grid = (int64_t)L.ntm * L.ntn * g.nz * g.ksplit;
  return L;
}

}  // namespace

int64_t gemm_i8_grid_items(const I8GemmArgs& g) {
  return i8_launch(g).grid * 256;
}

void gemm_sliced_i8(const I8GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  GAIN_CHECK_ARG(g.cfg >= 0 && g.cfg < kI8Configs, "int8 gemm: config %d",
                 g.cfg);
  GAIN_CHECK_ARG(g.Kp > 0 && g.Kp % 32 == 0, "int8 gemm: Kp %lld",
                 (long long)g.Kp);
  GAIN_CHECK_ARG(g.ksplit >= 1 && g.nz >= 1 && g.row0 >= 0,
                 "int8 gemm: bad split / z / row0");
  const I8Launch L = i8_launch(g);
  // every split's K range must keep the int32 level sums exact
  GAIN_CHECK_ARG((int64_t)L.kb_per * 32 <= i8_max_krange(g.S),
                 "int8 gemm: K range %lld per split exceeds %lld for %d "
                 "slices (use more splits)",
                 (long long)L.kb_per * 32, (long long)i8_max_krange(g.S),
                 g.S);
  GAIN_CHECK_ARG(L.grid < (1ll << 31), "int8 gemm: grid too large");
  const int key = g.S * 100 + g.cfg;
  const dim3 grid((unsigned)L.grid), block(256);
  switch (key) {
#define GAIN_I8_CASE(SS, CC)                                                \
  case SS * 100 + CC:                                                       \
    hipLaunchKernelGGL((k_gemm_i8<SS, kI8Cfg[CC].tm, kI8Cfg[CC].tn,         \
                                  kI8Cfg[CC].wm, kI8Cfg[CC].wn>),           \
                       grid, block, 0, s, g, L.ntm, L.ntn, L.rb0, L.kb_per); \
    break;
#define GAIN_I8_S(SS) \
  GAIN_I8_CASE(SS, 0) \
  GAIN_I8_CASE(SS, 1) \
  GAIN_I8_CASE(SS, 2) \
  GAIN_I8_CASE(SS, 3) \
  GAIN_I8_CASE(SS, 4) \
  GAIN_I8_CASE(SS, 5)
    GAIN_I8_S(5)
    GAIN_I8_S(6)
    GAIN_I8_S(7)
#undef GAIN_I8_S
#undef GAIN_I8_CASE
    default:
      fail(GAIN_ERR_ARG, "int8 slice count %d not supported (5..7)", g.S);
  }
  check_launch("gemm_sliced_i8");
}

}  // namespace gain
