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

#include <cmath>
#include <cstdint>

#include "common.hpp"
#include "gemm.hpp"
#include "kernels.hpp"
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
// Tile configurations: wave tile (32 TM) x (32 TN), WM x WN waves; `lds`:
// operands staged through LDS (k_gemm_i8_lds / k_gemm_i8_res_lds), else read
// per wave.
struct I8Cfg {
  int tm, tn, wm, wn, lds;
};
constexpr I8Cfg kI8Cfg[kI8Configs] = {
    {1, 2, 2, 2, 0},  // 64 x 128
    {2, 1, 2, 2, 0},  // 128 x 64
    {1, 2, 4, 1, 0},  // 128 x 64
    {1, 2, 1, 4, 0},  // 32 x 256
    {2, 1, 4, 1, 0},  // 256 x 32
    {2, 1, 1, 4, 0},  // 64 x 128
    {1, 2, 2, 2, 1},  // 64 x 128 through LDS
    {2, 1, 2, 2, 1},  // 128 x 64 through LDS
    // 8 waves (two per SIMD) of 32 x 32
    {1, 1, 4, 2, 0},  // 128 x 64
    {1, 1, 2, 4, 0},  // 64 x 128
    {1, 1, 4, 2, 1},  // 128 x 64 through LDS
    {1, 1, 2, 4, 1},  // 64 x 128 through LDS
};

const char* i8_config_name(int cfg) {
  static const char* names[kI8Configs] = {
      "64x128/w32x64", "128x64/w64x32", "128x64/w32x64", "32x256/w32x64",
      "256x32/w64x32", "64x128/w64x32", "64x128/w32x64/lds",
      "128x64/w64x32/lds", "128x64/8w", "64x128/8w", "128x64/8w/lds",
      "64x128/8w/lds"};
  return cfg >= 0 && cfg < kI8Configs ? names[cfg] : "?";
}

bool i8_config_ok(int cfg, int S) {
  return cfg >= 0 && cfg < kI8Configs && S >= 5 && S <= 7;
}

// Workgroup tile -> (z, split, tile row, tile column): XCD-aware remap of
// the linear id, splits outside tiles, grouped tile order.
__device__ inline void i8_tile(int ntm, int ntn, int ksplit, int* z, int* ksi,
                               int* tm, int* tn) {
  const int ntiles = ntm * ntn;
  const int per_z = ntiles * ksplit;
  const int p = xcd_remap(blockIdx.x, gridDim.x);
  *z = p / per_z;
  const int rem = p - *z * per_z;
  *ksi = rem / ntiles;
  // with a few column tiles, the column tiles of one row tile run back to
  // back (same XCD, same time), so the second one's reads of the shared A
  // blocks hit L2
  tile_order(rem - *ksi * ntiles, ntm, ntn, ntn <= 4 ? 1 : 8, tm, tn);
}

template <int S, int TM, int TN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_i8(I8GemmArgs g,
                                                          int ntm, int ntn,
                                                          int rb0,
                                                          int kb_per) {
  constexpr int BMB = TM * WM, BNB = TN * WN;  // tile size in 32-row blocks
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
  int z, ksi, tm, tn;
  i8_tile(ntm, ntn, g.ksplit, &z, &ksi, &tm, &tn);
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

  // PD register buffers: block kb + PD - 1 is requested before the
  // products of block kb, so each load has PD - 1 blocks of products to
  // land (the final records stream from HBM in the projection)
  // (two waves per SIMD with 8-wave workgroups: fewer buffers fit)
  constexpr int PD = (S <= 6 && WM * WN == 4) ? 3 : 2;
  i4v fa[PD][S][TM], fb[PD][S][TN];
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

#pragma unroll
  for (int q = 0; q < PD - 1; ++q)
    if (kb0 + q < kb1) load(fa[q], fb[q], kb0 + q);
  int kb = kb0;
  for (; kb + PD - 1 < kb1; kb += PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) {
      // buffer (q + PD - 1) % PD was consumed by the previous step
      if (kb + q + PD - 1 < kb1)
        load(fa[(q + PD - 1) % PD], fb[(q + PD - 1) % PD], kb + q + PD - 1);
      mma(fa[q], fb[q]);
    }
  }
  // fewer than PD blocks remain, already in buffers 0, 1, ...
#pragma unroll
  for (int q = 0; q < PD - 1; ++q)
    if (kb + q < kb1) mma(fa[q], fb[q]);

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

// Chunk swizzle inside a 1 KB block: 16-byte chunk c of the block is kept
// in LDS slot sigma(c) (an involution: bit 0 flipped by bits 3 and 4).  The
// fragment read of lane l takes chunk 2 (l % 32) + l / 32; with the swizzle
// the slots of every 8 consecutive lanes are distinct modulo 8 and those of
// every 16 distinct modulo 16, and the 128-bit fragment reads run without
// bank conflicts (SQ_LDS_BANK_CONFLICT 0, against 4 conflict cycles per
// read when bit 0 is flipped by bits 4 and 5).
__device__ inline int i8_sigma(int c) {
  return c ^ (((c >> 3) ^ (c >> 4)) & 1);
}

template <int N>
__device__ inline void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  // vmcnt in bits [3:0] and [15:14]; expcnt and lgkmcnt left at "no wait"
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) |
                             (15 << 8));
}

// Wait until at most `later` * DPW of this wave's loads are in flight
// (`later` stages issued after the one needed), for later <= MAXL.
template <int DPW, int MAXL>
__device__ inline void wait_stages(int later) {
  if constexpr (MAXL > 0) {
    if (later >= MAXL) {
      wait_vmcnt<MAXL * DPW>();
      return;
    }
    wait_stages<DPW, MAXL - 1>(later);
  } else {
    wait_vmcnt<0>();
  }
}

// Operands staged through LDS: workgroup tile (32 TM WM) x (32 TN WN), 4
// waves of (32 TM) x (32 TN) (the S level accumulators of a wave stay
// within the 256 accumulation registers), NST k-block stages filled by
// direct global -> LDS loads (global_load_lds_dwordx4: each wave-wide load
// copies one 1 KB block, lane l placing 16 bytes at slot l, so lane l
// fetches chunk sigma(l)).  Stage kb + NST - 1 is requested before the
// products of stage kb, so the loads have NST - 1 k blocks of products to
// land and use no registers; one barrier per k block (after it every wave
// has finished with stage kb - 1, whose buffer is refilled next, and every
// wave's loads of stage kb have landed).  A fragments of all slices are
// held per stage, B fragments are read once per slice.
template <int S, int TM, int TN, int WM, int WN, int NST>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_i8_lds(
    I8GemmArgs g, int ntm, int ntn, int rb0, int kb_per) {
  constexpr int NW = WM * WN;                  // waves
  constexpr int BMB = TM * WM, BNB = TN * WN;  // 32-row blocks per side
  constexpr int NBLK = S * (BMB + BNB);        // 1 KB blocks per stage
  // padded to a multiple of NW so every wave issues the same number of
  // loads (the padding slots repeat block 0 and are never read)
  constexpr int NBLKP = (NBLK + NW - 1) / NW * NW;
  constexpr int DPW = NBLKP / NW;              // block loads per wave
  static_assert(S * TM * TN * 16 <= 256, "accumulators beyond 256 regs");
  __shared__ __attribute__((aligned(1024))) int8_t lds[NST * NBLKP * 1024];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave % WM, wn = wave / WM;
  int z, ksi, tm, tn;
  i8_tile(ntm, ntn, g.ksplit, &z, &ksi, &tm, &tn);
  const int64_t nkb = g.Kp >> 5;
  const int kb0 = ksi * kb_per;
  const int kb1 = (int)min((int64_t)kb0 + kb_per, nkb);
  const int arb0 = rb0 + tm * BMB, brb0 = tn * BNB;
  const int src_off = 16 * i8_sigma(lane);
  const int8_t* Az = g.A + z * g.a_z;

  // block q = wave + NW t of a stage: A blocks first (q = BMB i + rb:
  // slice i, row block rb of the tile), then B blocks
  auto issue = [&](int stage, int kb) {
    int8_t* dst = lds + stage * (NBLKP * 1024);
#pragma unroll
    for (int t = 0; t < DPW; ++t) {
      const int q = wave + NW * t;
      const int8_t* src;
      if (q < S * BMB || q >= NBLK) {
        const int qa = q < NBLK ? q : 0;
        const int i = qa / BMB, rb = qa % BMB;
        src = Az + i * g.a_plane + ((int64_t)(arb0 + rb) * nkb + kb) * 1024;
      } else {
        const int i = (q - S * BMB) / BNB, rb = (q - S * BMB) % BNB;
        src = g.B + i * g.b_plane + ((int64_t)(brb0 + rb) * nkb + kb) * 1024;
      }
      __builtin_amdgcn_global_load_lds(
          (const void*)(src + src_off),
          (__attribute__((address_space(3))) void*)(dst + q * 1024), 16, 0,
          0);
    }
  };
  // request t of those issue() makes
  auto issue_one = [&](int stage, int kb, int t) {
    int8_t* dst = lds + stage * (NBLKP * 1024);
    const int q = wave + NW * t;
    const int8_t* src;
    if (q < S * BMB || q >= NBLK) {
      const int qa = q < NBLK ? q : 0;
      const int i = qa / BMB, rb = qa % BMB;
      src = Az + i * g.a_plane + ((int64_t)(arb0 + rb) * nkb + kb) * 1024;
    } else {
      const int i = (q - S * BMB) / BNB, rb = (q - S * BMB) % BNB;
      src = g.B + i * g.b_plane + ((int64_t)(brb0 + rb) * nkb + kb) * 1024;
    }
    __builtin_amdgcn_global_load_lds(
        (const void*)(src + src_off),
        (__attribute__((address_space(3))) void*)(dst + q * 1024), 16, 0, 0);
  };

  i16v acc[S][TM][TN];
#pragma unroll
  for (int L = 0; L < S; ++L)
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[L][a][b][e] = 0;

  const int rd_off = 16 * i8_sigma(2 * (lane & 31) + (lane >> 5));
#pragma unroll
  for (int st = 0; st < NST - 1; ++st)
    if (kb0 + st < kb1) issue(st, kb0 + st);
  int stage = 0;
  int kb = kb0;
  // k blocks that request a stage, in a loop of their own: the fragment
  // reads first, then the requests spread over the first products, one
  // after each (as in k_gemm_i8_res_lds; the projection 3-4% faster than
  // with the requests issued between the barrier and the reads)
  for (; kb + NST - 1 < kb1; ++kb) {
    wait_vmcnt<(NST >= 3 ? (NST - 2) * DPW : 0)>();
    __builtin_amdgcn_s_barrier();
    const int slot = stage == 0 ? NST - 1 : stage - 1;
    const int8_t* base = lds + stage * (NBLKP * 1024) + rd_off;
    i4v fa[S][TM];
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
      for (int a = 0; a < TM; ++a)
        fa[i][a] = *(const i4v*)(base + (BMB * i + TM * wm + a) * 1024);
#pragma unroll
    for (int j = 0; j < S; ++j) {
      i4v fb[TN];
#pragma unroll
      for (int b = 0; b < TN; ++b)
        fb[b] = *(const i4v*)(base + (S * BMB + BNB * j + TN * wn + b) * 1024);
#pragma unroll
      for (int i = 0; i < S - j; ++i)
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < TN; ++b) {
            acc[i + j][a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                fa[i][a], fb[b], acc[i + j][a][b], 0, 0, 0);
            const int m = (i * TM + a) * TN + b;
            if (j == 0 && m < DPW) {
              __builtin_amdgcn_sched_barrier(0);
              issue_one(slot, kb + NST - 1, m);
              __builtin_amdgcn_sched_barrier(0);
            }
          }
    }
    stage = stage == NST - 1 ? 0 : stage + 1;
  }
  for (; kb < kb1; ++kb) {
    // loads of later stages may stay in flight: min(NST - 2, left) of them
    const int later = min(NST - 2, kb1 - 1 - kb);
    if (later >= 2)
      wait_vmcnt<(NST >= 4 ? 2 * DPW : 0)>();
    else if (later == 1)
      wait_vmcnt<(NST >= 3 ? DPW : 0)>();
    else
      wait_vmcnt<0>();
    static_assert(NST <= 4, "wait counts cover up to 2 later stages");
    __builtin_amdgcn_s_barrier();
    if (kb + NST - 1 < kb1) issue(stage == 0 ? NST - 1 : stage - 1,
                                  kb + NST - 1);
    const int8_t* base = lds + stage * (NBLKP * 1024) + rd_off;
    i4v fa[S][TM];
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
      for (int a = 0; a < TM; ++a)
        fa[i][a] = *(const i4v*)(base + (BMB * i + TM * wm + a) * 1024);
#pragma unroll
    for (int j = 0; j < S; ++j) {
      i4v fb[TN];
#pragma unroll
      for (int b = 0; b < TN; ++b)
        fb[b] = *(const i4v*)(base + (S * BMB + BNB * j + TN * wn + b) * 1024);
#pragma unroll
      for (int i = 0; i < S - j; ++i)
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
          for (int b = 0; b < TN; ++b)
            acc[i + j][a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                fa[i][a], fb[b], acc[i + j][a][b], 0, 0, 0);
    }
    stage = stage == NST - 1 ? 0 : stage + 1;
  }

  double* C = g.C + z * g.c_z + ksi * g.c_ks;
  const int* eA = g.eA + z * g.ea_z;
  const int arb = arb0 + TM * wm, brb = brb0 + TN * wn;
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
  return i8_launch(g).grid * 64 * kI8Cfg[g.cfg].wm * kI8Cfg[g.cfg].wn;
}

void gemm_sliced_i8(const I8GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  GAIN_CHECK_ARG(i8_config_ok(g.cfg, g.S),
                 "int8 gemm: configuration %d with %d slices", g.cfg, g.S);
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
  const dim3 grid((unsigned)L.grid),
      block(64 * kI8Cfg[g.cfg].wm * kI8Cfg[g.cfg].wn);
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
  GAIN_I8_CASE(SS, 5) \
  GAIN_I8_CASE(SS, 8) \
  GAIN_I8_CASE(SS, 9)
    GAIN_I8_S(5)
    GAIN_I8_S(6)
    GAIN_I8_S(7)
#define GAIN_I8_LDS(SS, CC, NS)                                             \
  case SS * 100 + CC:                                                       \
    hipLaunchKernelGGL((k_gemm_i8_lds<SS, kI8Cfg[CC].tm, kI8Cfg[CC].tn,     \
                                      kI8Cfg[CC].wm, kI8Cfg[CC].wn, NS>),   \
                       grid, block, 0, s, g, L.ntm, L.ntn, L.rb0, L.kb_per); \
    break;
    GAIN_I8_LDS(5, 6, 4)
    GAIN_I8_LDS(5, 7, 4)
    GAIN_I8_LDS(6, 6, 4)
    GAIN_I8_LDS(6, 7, 4)
    GAIN_I8_LDS(7, 6, 3)
    GAIN_I8_LDS(7, 7, 3)
    GAIN_I8_LDS(5, 10, 2)
    GAIN_I8_LDS(5, 11, 2)
    GAIN_I8_LDS(6, 10, 2)
    GAIN_I8_LDS(6, 11, 2)
    GAIN_I8_LDS(7, 10, 2)
    GAIN_I8_LDS(7, 11, 2)
#undef GAIN_I8_LDS
#undef GAIN_I8_S
#undef GAIN_I8_CASE
    default:
      fail(GAIN_ERR_ARG, "int8 gemm: %d slices with configuration %s not "
           "supported", g.S, i8_config_name(g.cfg));
  }
  check_launch("gemm_sliced_i8");
}

// ---------------------------------------------------------------------------
// CRT form.  Moduli: odd, pairwise coprime, <= 255.
namespace {

constexpr int kCrtModuli[kCrtMaxModuli] = {255, 253, 251, 247, 241, 239, 233,
                                           229, 227, 223, 217, 211, 199, 197};

int64_t inv_mod(int64_t a, int64_t m) {  // a^{-1} mod m, gcd(a, m) = 1
  int64_t t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
  while (nr != 0) {
    const int64_t q = r / nr;
    const int64_t tt = t - q * nt;
    t = nt;
    nt = tt;
    const int64_t rr = r - q * nr;
    r = nr;
    nr = rr;
  }
  return ((t % m) + m) % m;
}

// a mod m in [-(m - 1) / 2, (m - 1) / 2] (m odd)
int64_t sym_res(int64_t a, int64_t m) {
  int64_t r = ((a % m) + m) % m;
  return r > (m - 1) / 2 ? r - m : r;
}

CrtConsts make_crt(int nm) {
  CrtConsts c{};
  c.nm = nm;
  for (int i = 0; i < nm; ++i) {
    const int64_t m = kCrtModuli[i];
    c.m[i] = (int)m;
    c.minv[i] = 1.0f / (float)m;
    c.minv_d[i] = 1.0 / (double)m;
  }
  // explicit form: C_i = M_i (M_i^{-1} mod m_i), split at 2^es with
  // es = bit length of M - 40
  typedef unsigned __int128 u128;
  u128 M = 1;
  for (int i = 0; i < nm; ++i) M *= (u128)kCrtModuli[i];
  int bits = 0;
  for (u128 t = M; t; t >>= 1) ++bits;
  const int es = std::max(0, bits - 40);
  auto split = [&](u128 v, double* hi, double* lo) {
    const u128 h = (v >> es) << es;
    *hi = std::ldexp((double)(uint64_t)(v >> es), es);  // < 2^40: exact
    *lo = (double)(v - h);                               // < 2^es, rounded
  };
  for (int i = 0; i < nm; ++i) {
    const int64_t m = kCrtModuli[i];
    const u128 Mi = M / (u128)m;
    const int64_t inv = inv_mod((int64_t)(Mi % (u128)m), m);
    const u128 C = Mi * (u128)inv;  // < M
    // C / M rounded to FP64 from a 64-bit-mantissa quotient
    c.ef[i] = (double)((long double)C / (long double)M);
    split(C, &c.ech[i], &c.ecl[i]);
  }
  split(M, &c.emh, &c.eml);
  // pairs of moduli, then the Garner weights over the pair moduli
  c.np = (nm + 1) / 2;
  int64_t P[CrtConsts::kMaxPairs];
  for (int k = 0; k < c.np; ++k) {
    const int64_t a = kCrtModuli[2 * k];
    if (2 * k + 1 < nm) {
      const int64_t b = kCrtModuli[2 * k + 1];
      c.pinv[k] = (float)sym_res(inv_mod(a, b), b);
      P[k] = a * b;
    } else {
      c.pinv[k] = 0.0f;
      P[k] = a;
    }
    c.P[k] = (double)P[k];
    c.Pinv[k] = 1.0 / (double)P[k];
  }
  for (int k = 0; k < c.np; ++k) {
    int64_t pw[CrtConsts::kMaxPairs];  // P_0 ... P_{j-1} mod P_k
    int64_t prod = 1;
    for (int j = 0; j < k; ++j) {
      pw[j] = prod;
      prod = prod * P[j] % P[k];
    }
    const int64_t inv = k == 0 ? 1 : inv_mod(prod, P[k]);
    c.pgi[k] = (double)sym_res(inv, P[k]);
    for (int j = 0; j < k; ++j)
      c.pnw[k][j] = (double)sym_res(-pw[j] * inv % P[k], P[k]);
  }
  return c;
}

}  // namespace

const CrtConsts& crt_consts(int nm) {
  static const CrtConsts c11 = make_crt(11), c12 = make_crt(12),
                         c14 = make_crt(14);
  if (nm == 11) return c11;
  if (nm == 12) return c12;
  if (nm == 14) return c14;
  fail(GAIN_ERR_ARG, "CRT moduli count %d not supported (11, 12, 14)", nm);
}

// Operand width of a CRT product of length K and the row scaling it goes
// with.  Rows scaled to a 2-norm below 2^{beta-1} and rounded entry by
// entry (each entry moves by at most 1/2) have norms below
// 2^{beta-1} + sqrt(K)/2, so by Cauchy-Schwarz every product
// sum_k X_A X_B stays below (2^{beta-1} + sqrt(K)/2)^2; rows scaled to a
// maximum below 2^{beta-1} keep it below K 4^{beta-1}.  Either bound must
// stay below P (1/2 - 3.5e-7), so that |X| / M < 1/2 - 3e-7 and the
// explicit reconstruction's quotient estimate is never ambiguous.  The
// norm scaling gives about log2(K)/2 more bits per operand and is used
// whenever its width fits the 50-bit cap (at most 50-bit integers: their
// residues are taken exactly in FP64, sym_mod_d); past the cap the scaling
// by the maximum keeps more bits below each row's largest entry.
static int crt_width(int nm, int64_t K, bool* norm) {
  GAIN_CHECK_ARG(K > 0 && nm >= 1 && nm <= kCrtMaxModuli,
                 "CRT: %d moduli, product length %lld", nm, (long long)K);
  long double P = 1.0L;
  for (int i = 0; i < nm; ++i) P *= (long double)kCrtModuli[i];
  const long double ln =
      sqrtl(P * (0.5L - 3.5e-7L)) - sqrtl((long double)K) / 2.0L;
  if (ln > 256.0L) {
    int b = (int)floorl(log2l(ln));
    if (ldexpl(1.0L, b) >= ln) --b;  // strict inequality
    if (b + 1 <= 50) {
      *norm = true;
      return b + 1;
    }
  }
  // 2 K 4^{beta-1} < P 2^{-1e-6} (no K that is a multiple of 64 below 2e5
  // comes within that margin of a boundary)
  const long double lm =
      (log2l(P) - 1.0L - log2l((long double)K) - 1e-6L) / 2.0L;
  int b = (int)floorl(lm);
  if ((long double)b >= lm) --b;  // strict inequality
  GAIN_CHECK_ARG(b >= 8, "CRT: no integer width for K = %lld", (long long)K);
  *norm = false;
  return std::min(b, 49) + 1;
}

int crt_beta(int nm, int64_t K) {
  bool norm;
  return crt_width(nm, K, &norm);
}

bool crt_norm_scaled(int nm, int64_t K) {
  bool norm;
  crt_width(nm, K, &norm);
  return norm;
}

// x mod m in [-(m - 1) / 2, (m - 1) / 2] for an odd modulus m and an
// integer-valued x with |x| < 2^51, minv = fl(1/m).  x * minv is within
// |x/m| 2^-52 of x/m, and x/m (m odd) is at least 1/(2m) away from every
// half-integer, so rint(x * minv) = rint(x/m) and the result is exact.
__host__ __device__ inline double sym_mod_d(double x, double m,
                                            double minv) {
  return fma(-rint(x * minv), m, x);
}

// Same in FP32 for an integer-valued |x| < 2^22 (error |x/m| 2^-23 < 1/(2m)).
__host__ __device__ inline float sym_mod_f(float x, float m, float minv) {
  return fmaf(-rintf(x * minv), m, x);
}

// Exponent e of every padded row (0 for a zero or padding row): with
// `norm`, a bound of the 2-norm, ||x||_2 < 2^e (hence |x_k| < 2^e), else of
// the maximum, |x_k| < 2^e.  One workgroup per row; the sum of squares runs
// in a fixed order, so a row gets the same exponent on every run.  The
// computed sum is within K 2^-53 of the exact one (K < 2^30), so the square
// root of the sum times 1 + 2^-30, times 1 + 2^-40, bounds the norm from
// above.  Rows whose squares leave the FP64 range (entries beyond 2^500 or
// all below 2^-500) take the bound sqrt(K) max|x_k| instead.
// Exponent e of a row from its largest magnitude mx and its sum of squares
// ss (any summation order: the bound carries a 2^-30 margin): with norm,
// ||x|| < 2^e; without, max |x_k| < 2^e; 0 for a zero row.
__device__ inline int crt_row_expo(double mx, double ss, bool norm, int K) {
  int e = 0;
  if (mx > 0.0) {
    const double nb = sqrt(ss * (1.0 + 0x1p-30)) * (1.0 + 0x1p-40);
    if (!norm) {
      frexp(mx, &e);  // |x_k| < 2^e
    } else if (ss > 0x1p-1000 && nb < 0x1p500) {
      frexp(nb, &e);  // ||x|| <= nb < 2^e
    } else {
      frexp(mx, &e);  // |x_k| < 2^e, ||x|| < sqrt(K) 2^e
      int lk = 0;
      while ((int64_t)1 << lk < K) ++lk;
      e += (lk + 1) / 2;
    }
  }
  return e;
}

__global__ __launch_bounds__(256) void k_row_expo(const double* __restrict__ src,
                                                  int64_t ld, int rows, int K,
                                                  bool norm,
                                                  int* __restrict__ expo) {
  __shared__ double red[2][4];
  const int r = blockIdx.x;
  const int tid = threadIdx.x;
  double mx = 0.0, ss = 0.0;
  if (r < rows) {
    const double* x = src + (int64_t)r * ld;
    double m2 = 0.0, s2 = 0.0;
    int k = tid;
    for (; k + 256 < K; k += 512) {
      const double a = x[k], b = x[k + 256];
      mx = fmax(mx, fabs(a));
      m2 = fmax(m2, fabs(b));
      ss = fma(a, a, ss);
      s2 = fma(b, b, s2);
    }
    if (k < K) {
      const double a = x[k];
      mx = fmax(mx, fabs(a));
      ss = fma(a, a, ss);
    }
    mx = fmax(mx, m2);
    ss += s2;
  }
  // butterfly sums: every lane adds the same pairs (a + b = b + a exactly)
  for (int off = 32; off >= 1; off >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, off));
    ss += __shfl_xor(ss, off);
  }
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = mx;
    red[1][tid >> 6] = ss;
  }
  __syncthreads();
  if (tid == 0) {
    mx = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
    ss = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    expo[r] = crt_row_expo(mx, ss, norm, K);
  }
}

// Residues of the beta-bit integers rint(x 2^{beta-1-e}) of a tile row
// (32 rows) over kKch k blocks: thread t converts row t / 8, columns
// 4 (t % 8) .. + 3 of each 32-column block, and every wave writes 256
// contiguous bytes of each residue plane (the workgroup one whole 1 KB
// block).  Rows past `rows` (padding) and columns past K are zero.
constexpr int kSliceKch = 16;  // k blocks per workgroup
template <int NM>
__global__ __launch_bounds__(256) void k_slice_crt(
    const double* __restrict__ src, int64_t ld, int rows, int K, int64_t Kp,
    int64_t plane, int8_t* __restrict__ planes, int* __restrict__ expo,
    int beta, CrtConsts cc) {
  const int t = threadIdx.x;
  const int rr = t >> 3, kk = 4 * (t & 7);
  const int r = blockIdx.x * 32 + rr;
  const bool live = r < rows;
  const int e = live ? expo[r] : 0;
  const int sh = live ? beta - 1 - e : 0;
  const double* x = src + (live ? (int64_t)r * ld : 0);
  const int64_t nkb = Kp >> 5;
  const int64_t kb0 = (int64_t)blockIdx.y * kSliceKch;
  const int64_t kb1 = min(kb0 + kSliceKch, nkb);
  int8_t* tile = planes + (int64_t)blockIdx.x * (Kp << 5) + rr * 32 + kk;
  for (int64_t kb = kb0; kb < kb1; ++kb) {
    const int64_t k0 = kb * 32 + kk;
    double xv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      xv[q] = (live && k0 + q < K) ? x[k0 + q] : 0.0;
    uint32_t w[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) w[i] = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // |X| <= 2^{beta-1} <= 2^49: exact in double
      const double X = rint(ldexp(xv[q], sh));
#pragma unroll
      for (int i = 0; i < NM; ++i) {
        const double v = sym_mod_d(X, (double)cc.m[i], cc.minv_d[i]);
        w[i] |= (uint32_t)(uint8_t)(int8_t)(int)v << (8 * q);
      }
    }
    int8_t* dst = tile + kb * 1024;
#pragma unroll
    for (int i = 0; i < NM; ++i) *(uint32_t*)(dst + i * plane) = w[i];
  }
}

void slice_rows_crt(int nm, int beta, const double* src, int64_t ld,
                    int rows, int K, int8_t* planes, int64_t plane,
                    int* expo, hipStream_t s) {
  const int64_t Rp = i8_rows_pad(rows), Kp = i8_kpad(K);
  GAIN_CHECK_ARG(rows >= 0 && K > 0 && plane >= Rp * Kp,
                 "slice_rows_crt: plane of %lld bytes < %lld x %lld",
                 (long long)plane, (long long)Rp, (long long)Kp);
  GAIN_CHECK_ARG(((uintptr_t)planes & 3) == 0 && (plane & 3) == 0,
                 "slice_rows_crt: planes must be 4-byte aligned");
  GAIN_CHECK_ARG(beta >= 9 && beta <= 50, "slice_rows_crt: beta %d", beta);
  GAIN_CHECK_ARG(beta == crt_beta(nm, Kp),
                 "slice_rows_crt: width %d is not that of %d moduli at K %lld",
                 beta, nm, (long long)Kp);
  const CrtConsts& cc = crt_consts(nm);
  hipLaunchKernelGGL(k_row_expo, dim3((unsigned)Rp), dim3(256), 0, s, src,
                     ld, rows, K, crt_norm_scaled(nm, Kp), expo);
  check_launch("row_expo");
  const dim3 grid((unsigned)(Rp / 32),
                  (unsigned)((Kp / 32 + kSliceKch - 1) / kSliceKch));
  switch (nm) {
#define GAIN_CRT_SLICE(NN)                                                   \
  case NN:                                                                   \
    hipLaunchKernelGGL(k_slice_crt<NN>, grid, dim3(256), 0, s, src, ld,      \
                       rows, K, Kp, plane, planes, expo, beta, cc);          \
    break;
    GAIN_CRT_SLICE(11)
    GAIN_CRT_SLICE(12)
    GAIN_CRT_SLICE(14)
#undef GAIN_CRT_SLICE
    default:
      fail(GAIN_ERR_ARG, "CRT moduli count %d not supported", nm);
  }
  check_launch("slice_rows_crt");
}

// ---------------------------------------------------------------------------
// The half line strength's X built and cut into residue planes in one pass
// (no FP64 X in HBM).  k_slot_desc flattens each (column, v) slot of the
// initial basis once per call: for a slot of one primitive its record
// offset t0, term count, repres offset (lambda sdeg Ntot + t0) and the
// initial-side sign (dipole_GPU.cu:437-438) in w; for a slot of several the
// CSR range with w = 2; an empty slot has no terms.
__global__ __launch_bounds__(256) void k_slot_desc(ScatterArgs a,
                                                   int4* __restrict__ desc) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= a.nslot) return;
  const int q0 = a.slot_ptr[s], q1 = a.slot_ptr[s + 1];
  int4 d = make_int4(0, 0, 0, 0);
  if (q1 - q0 == 1) {
    const int r = a.slot_prim[q0];
    const int icase = a.icontr[r];
    const int lam = a.icontr[r + a.nprim];
    const int t0 = a.ijt[icase];
    const int K = a.kval[r], t = a.tau[r];
    d = make_int4(t0, a.N[icase], lam * a.sdeg * a.ntot + t0,
                  ((K % 3) * t + K) & 1);
  } else if (q1 > q0) {
    d = make_int4(q0, q1, 0, 2);
  }
  desc[s] = d;
}

struct XSliceArgs {
  ScatterArgs sa;  // basis, records and vectors (its X is not used)
  const int4* desc;
  int ncol, nc, rows;
  int64_t Kp, plane;
  int8_t* planes;
  int* expo;
  int beta;
  bool norm;
};

// A slot of several primitives: the scatter's loop (scatter_body).
__device__ __noinline__ double x_value_multi(const ScatterArgs& a,
                                             const double* cv, int64_t dof,
                                             int q0, int q1) {
  double acc = 0.0;
  for (int q = q0; q < q1; ++q) {
    const int r = a.slot_prim[q];
    const int icase = a.icontr[r];
    const int lam = a.icontr[r + a.nprim];
    const int t0 = a.ijt[icase];
    const int ne = a.N[icase];
    const double* c = cv + t0;
    const double* p = a.repres + (int64_t)lam * a.sdeg * a.ntot + t0 + dof;
    const int K = a.kval[r], t = a.tau[r];
    double x = 0.0;
    for (int k = 0; k < ne; ++k) x = fma(c[k], p[k], x);
    acc += (((K % 3) * t + K) & 1) ? -x : x;
  }
  return acc;
}

// X[vec][col][v] with the scatter's arithmetic (identical values).
__device__ inline double x_value(const XSliceArgs& a, const double* cv,
                                 int64_t dof, int s) {
  const int4 d = a.desc[s];
  if (d.w == 2) return x_value_multi(a.sa, cv, dof, d.x, d.y);
  const double* c = cv + d.x;
  const double* p = a.sa.repres + dof + d.z;
  double x = 0.0;
  for (int k = 0; k < d.y; ++k) x = fma(c[k], p[k], x);
  return 0.0 + (d.w ? -x : x);
}

// One workgroup: 4 consecutive X rows (vector, column) of R * 8 k blocks.
// Thread t holds row (t / 8) % 4, columns 4 (t % 8) .. + 3 of k blocks
// 8 j + t / 32 (j < R) in registers; the row's norm comes from butterflies
// over the lanes of that row and the four waves' partials in a fixed
// order, then the registers are cut into residues: each 32 lanes write the
// four rows' 128 contiguous bytes of a 1 KB block of every plane (the
// layout of k_slice_crt).  ONE: every slot holds at most one primitive of
// one term, and the values take no branch (an empty or padding slot loads
// entry 0 and is zeroed), so the loads of all of a thread's values can be
// in flight together.
template <int NM, int R, bool ONE>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(ONE ? 3 : 1))) void k_scatter_slice_crt(
    XSliceArgs a, CrtConsts cc) {
  __shared__ double red[2][4][4];  // [max | sum of squares][wave][row]
  const int t = threadIdx.x;
  const int q = t & 7, rr = (t >> 3) & 3, kbo = t >> 5;
  const int n = blockIdx.x * 4 + rr;
  const bool live = n < a.rows;
  const int vec = live ? n / a.ncol : 0;
  const int col = live ? n - vec * a.ncol : 0;
  const double* cv = a.sa.cbase + (live ? a.sa.c_off[vec] : 0);
  const int64_t dof = live ? (int64_t)a.sa.ideg[vec] * a.sa.ntot : 0;
  const int nkb = (int)(a.Kp >> 5);
  double x[R][4];
  double mx = 0.0, ss = 0.0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = (j * 8 + kbo) * 32 + 4 * q + u;
      const bool ok = live && k < a.nc;
      double v;
      if constexpr (ONE) {
        const int4 d = a.desc[ok ? col * a.nc + k : 0];
        const double xv = fma(cv[d.x], a.sa.repres[dof + d.z], 0.0);
        v = (ok && d.y) ? 0.0 + (d.w ? -xv : xv) : 0.0;
      } else {
        v = ok ? x_value(a, cv, dof, col * a.nc + k) : 0.0;
      }
      x[j][u] = v;
      mx = fmax(mx, fabs(v));
      ss = fma(v, v, ss);
    }
  }
  // lanes of one row in a wave differ in bits 0-2 and 5 of the lane id
#pragma unroll
  for (int off = 1; off <= 32; off <<= 1) {
    if (off == 8 || off == 16) continue;
    mx = fmax(mx, __shfl_xor(mx, off));
    ss += __shfl_xor(ss, off);
  }
  if ((t & 39) == 0) {  // q = 0 and lane bit 5 clear: one lane per row
    red[0][t >> 6][rr] = mx;
    red[1][t >> 6][rr] = ss;
  }
  __syncthreads();
  mx = fmax(fmax(red[0][0][rr], red[0][1][rr]),
            fmax(red[0][2][rr], red[0][3][rr]));
  ss = (red[1][0][rr] + red[1][1][rr]) + (red[1][2][rr] + red[1][3][rr]);
  const int e = crt_row_expo(mx, ss, a.norm, a.nc);
  if (q == 0 && kbo == 0) a.expo[n] = e;
  const int sh = a.beta - 1 - e;
  int8_t* tile = a.planes + (int64_t)(n >> 5) * (a.Kp << 5) + (n & 31) * 32 +
                 4 * q;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int kb = j * 8 + kbo;
    if (kb >= nkb) continue;
    uint32_t w[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) w[i] = 0u;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      // |X| <= 2^{beta-1} <= 2^49: exact in double
      const double X = rint(ldexp(x[j][u], sh));
#pragma unroll
      for (int i = 0; i < NM; ++i) {
        const double v = sym_mod_d(X, (double)cc.m[i], cc.minv_d[i]);
        w[i] |= (uint32_t)(uint8_t)(int8_t)(int)v << (8 * u);
      }
    }
    int8_t* dst = tile + (int64_t)kb * 1024;
#pragma unroll
    for (int i = 0; i < NM; ++i) *(uint32_t*)(dst + i * a.plane) = w[i];
  }
}

bool scatter_slice_crt_fits(int nm, int nc, int64_t repres_len) {
  return (nm == 11 || nm == 12 || nm == 14) && nc >= 1 &&
         i8_kpad(nc) <= 32 * 8 * 16 && repres_len < INT32_MAX;
}

void scatter_slice_rows_crt(int nm, int beta, const ScatterArgs& sa, int ncol,
                            int nc, int rows, int8_t* planes, int64_t plane,
                            int* expo, int4* desc, bool single,
                            hipStream_t s) {
  const int64_t Rp = i8_rows_pad(rows), Kp = i8_kpad(nc);
  GAIN_CHECK_ARG(rows >= 0 && ncol >= 1 && sa.nslot == ncol * nc &&
                     (int64_t)sa.nvec * ncol == rows,
                 "scatter_slice_rows_crt: %d rows of %d columns, %d vectors,"
                 " %d slots", rows, ncol, sa.nvec, sa.nslot);
  GAIN_CHECK_ARG(Kp <= 32 * 8 * 16,
                 "scatter_slice_rows_crt: K = %d beyond %d", nc, 32 * 8 * 16);
  GAIN_CHECK_ARG(plane >= Rp * Kp && ((uintptr_t)planes & 3) == 0 &&
                     (plane & 3) == 0,
                 "scatter_slice_rows_crt: plane of %lld bytes < %lld x %lld "
                 "or unaligned", (long long)plane, (long long)Rp,
                 (long long)Kp);
  GAIN_CHECK_ARG(beta == crt_beta(nm, Kp),
                 "scatter_slice_rows_crt: width %d is not that of %d moduli "
                 "at K %lld", beta, nm, (long long)Kp);
  hipLaunchKernelGGL(k_slot_desc, dim3((unsigned)((sa.nslot + 255) / 256)),
                     dim3(256), 0, s, sa, desc);
  check_launch("slot_desc");
  XSliceArgs a{};
  a.sa = sa;
  a.desc = desc;
  a.ncol = ncol;
  a.nc = nc;
  a.rows = rows;
  a.Kp = Kp;
  a.plane = plane;
  a.planes = planes;
  a.expo = expo;
  a.beta = beta;
  a.norm = crt_norm_scaled(nm, Kp);
  const CrtConsts& cc = crt_consts(nm);
  const dim3 grid((unsigned)(Rp / 4));
  const bool small = Kp <= 32 * 8 * 8;
  switch (nm) {
#define GAIN_XSLICE_R(NN, RR)                                                \
  if (single)                                                                \
    hipLaunchKernelGGL((k_scatter_slice_crt<NN, RR, true>), grid, dim3(256), \
                       0, s, a, cc);                                         \
  else                                                                       \
    hipLaunchKernelGGL((k_scatter_slice_crt<NN, RR, false>), grid,           \
                       dim3(256), 0, s, a, cc);
#define GAIN_XSLICE(NN)                                                      \
  case NN:                                                                   \
    if (small) {                                                             \
      GAIN_XSLICE_R(NN, 8)                                                   \
    } else {                                                                 \
      GAIN_XSLICE_R(NN, 16)                                                  \
    }                                                                        \
    break;
    GAIN_XSLICE(11)
    GAIN_XSLICE(12)
    GAIN_XSLICE(14)
#undef GAIN_XSLICE
#undef GAIN_XSLICE_R
    default:
      fail(GAIN_ERR_ARG, "CRT moduli count %d not supported", nm);
  }
  check_launch("scatter_slice_crt");
}

// Configurations of the per-modulus GEMM: wave tile (32 TM) x (32 TN), WM x
// WN waves.  One accumulator set per wave, so the wave tiles can be large
// (each residue plane enters one product only: the tile size sets the
// products per byte loaded).
constexpr I8Cfg kCrtCfg[kCrtConfigs] = {
    {4, 4, 2, 2, 0},  // 256 x 256, 4 waves of 128 x 128
    {2, 4, 2, 2, 0},  // 128 x 256, 4 waves of 64 x 128
    {4, 2, 2, 2, 0},  // 256 x 128, 4 waves of 128 x 64
    {2, 2, 2, 2, 0},  // 128 x 128, 4 waves of 64 x 64
    {2, 2, 2, 4, 0},  // 128 x 256, 8 waves of 64 x 64
    {4, 4, 2, 2, 1},  // 256 x 256 through LDS (4 stages)
    {2, 4, 2, 2, 1},  // 128 x 256 through LDS
    {2, 2, 2, 4, 1},  // 128 x 256, 8 waves, through LDS
    {4, 4, 2, 2, 4},  // 256 x 256 through LDS, 6 stages
    {4, 4, 2, 2, 2},  // 256 x 256, LDS + fragment prefetch, 4 stages
    {4, 4, 2, 2, 3},  // 256 x 256, LDS + fragment prefetch, 6 stages
    {2, 4, 2, 2, 3},  // 128 x 256, LDS + fragment prefetch, 6 stages
    {4, 4, 2, 2, 5},  // 256 x 256, prefetch, 4 stages of 2 k blocks
    {4, 2, 2, 4, 5},  // 256 x 256, 8 waves of 128 x 64, as above
};
// LDS pipeline of a configuration: stage count, fragment prefetch, k
// blocks per stage
constexpr int crt_nst(int lds) { return lds == 1 || lds == 2 || lds == 5 ? 4 : 6; }
constexpr bool crt_pf(int lds) { return lds == 2 || lds == 3 || lds == 5; }
constexpr int crt_kpb(int lds) { return lds == 5 ? 2 : 1; }

const char* crt_config_name(int cfg) {
  static const char* names[kCrtConfigs] = {
      "256x256/w128x128", "128x256/w64x128", "256x128/w128x64",
      "128x128/w64x64", "128x256/8w/w64x64", "256x256/w128x128/lds",
      "128x256/w64x128/lds", "128x256/8w/lds", "256x256/lds6",
      "256x256/lds4/pf", "256x256/lds6/pf", "128x256/lds6/pf",
      "256x256/lds4x2/pf", "256x256/8w/lds4x2/pf"};
  return cfg >= 0 && cfg < kCrtConfigs ? names[cfg] : "?";
}

// Residues of one 32 x 32 accumulator tile into the residue layout
// [row block][column][32 rows] (int8): lane l holds column l % 32, rows
// 8 w + 4 (l / 32) + (0..3) in word w; lanes l and l ^ 32 trade words so
// that each stores two runs of 8 rows (0-7 and 16-23 for l < 32, 8-15 and
// 24-31 otherwise), and every store instruction writes 32 whole 32-byte
// column segments.  `col` is the column's segment of the tile's row block.
__device__ inline void store_residue_tile(const i16v& acc, double m,
                                          double minv, int8_t* col,
                                          bool live, int lane) {
  uint32_t wd[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    uint32_t word = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double r = sym_mod_d((double)acc[4 * w + q], m, minv);
      word |= (uint32_t)(uint8_t)(int8_t)(int)r << (8 * q);
    }
    wd[w] = word;
  }
  const int hi = lane >> 5;
  const uint32_t o0 = __shfl_xor(hi ? wd[0] : wd[1], 32);
  const uint32_t o1 = __shfl_xor(hi ? wd[2] : wd[3], 32);
  if (!live) return;
  const uint2 s0 = hi ? make_uint2(o0, wd[1]) : make_uint2(wd[0], o0);
  const uint2 s1 = hi ? make_uint2(o1, wd[3]) : make_uint2(wd[2], o1);
  *(uint2*)(col + 8 * hi) = s0;
  *(uint2*)(col + 16 + 8 * hi) = s1;
}

// Workgroup tile of a CRT-form GEMM -> (modulus mi, operand matrix za,
// tile row tm, tile column tn).  Without A k-block ranges: as i8_tile, z =
// za * nm + mi outermost.  With ranges and several operand matrices (the
// three dipole components against one X): the modulus outermost, then
// groups of the nz matrices of one tile row walked column by column, so
// the nz tiles of a column run side by side on one XCD with the same k
// blocks (crt_tile_krange takes the union over the matrices) and share the
// B blocks in L2.
__device__ inline void crt_tile(const I8GemmArgs& g, int ntm, int ntn,
                                int* za, int* mi, int* tm, int* tn) {
  if (!g.a_kr || g.nz == 1) {
    int z, ksi;
    i8_tile(ntm, ntn, 1, &z, &ksi, tm, tn);
    *za = z / g.S;
    *mi = z - *za * g.S;
    return;
  }
  const int p = xcd_remap(blockIdx.x, gridDim.x);
  const int per_mod = ntm * g.nz * ntn;
  *mi = p / per_mod;
  int trow;
  tile_order(p - *mi * per_mod, ntm * g.nz, ntn, g.nz, &trow, tn);
  *tm = trow / g.nz;
  *za = trow - *tm * g.nz;
}

// k blocks [*k0, *k0 + *nk) that the workgroup tile with A row blocks
// [arb0, arb0 + nrb) multiplies: the union of the row blocks' nonzero
// ranges over every operand matrix (the nz matrices of one tile row then
// walk the same k blocks), widened to multiples of `align` (all nkb blocks
// without ranges; none for a zero tile).  The values are uniform over the
// workgroup.
__device__ inline void crt_tile_krange(const I8GemmArgs& g, int arb0,
                                       int nrb, int nkb, int align, int* k0,
                                       int* nk) {
  if (!g.a_kr) {
    *k0 = 0;
    *nk = nkb;
    return;
  }
  int lo = nkb, hi = 0;
  for (int z = 0; z < g.nz; ++z) {
    const int2* kr = g.a_kr + z * g.akr_z + arb0;
    for (int q = 0; q < nrb; ++q) {
      const int2 r = kr[q];
      if (r.x < r.y) {
        lo = min(lo, r.x);
        hi = max(hi, r.y);
      }
    }
  }
  if (lo >= hi) {
    *k0 = 0;
    *nk = 0;
    return;
  }
  lo = lo / align * align;
  hi = min(nkb, (hi + align - 1) / align * align);
  *k0 = __builtin_amdgcn_readfirstlane(lo);
  *nk = __builtin_amdgcn_readfirstlane(hi - lo);
}

// Per-modulus GEMM: workgroup z in [0, nz * nm) takes modulus z % nm of
// operand matrix z / nm; the int32 products of the residue planes are
// reduced modulo m and stored as int8 residues R[z][row block][n][32]
// (rows relative to the first row block; store_residue_tile).  With A
// k-block ranges only the tile's nonzero k blocks are multiplied.
template <int TM, int TN, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_i8_res(
    I8GemmArgs g, CrtConsts cc, int ntm, int ntn, int rb0, int nkb,
    int64_t ldr, int nr) {
  constexpr int BMB = TM * WM, BNB = TN * WN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave % WM, wn = wave / WM;
  int za, mi, tm, tn;
  crt_tile(g, ntm, ntn, &za, &mi, &tm, &tn);
  const int z = za * g.S + mi;
  int k0, nk;
  crt_tile_krange(g, rb0 + tm * BMB, BMB, nkb, 2, &k0, &nk);
  const int arb = rb0 + tm * BMB + wm * TM;
  const int brb = tn * BNB + wn * TN;
  const int lane_off = (lane & 31) * 32 + 16 * (lane >> 5);
  const int8_t* Ab = g.A + za * g.a_z + mi * g.a_plane +
                     ((int64_t)arb * nkb + k0) * 1024 + lane_off;
  const int8_t* Bb = g.B + mi * g.b_plane + ((int64_t)brb * nkb + k0) * 1024 +
                     lane_off;
  i16v acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0;
  i4v fa0[TM], fb0[TN], fa1[TM], fb1[TN];
  auto load = [&](i4v(&fa)[TM], i4v(&fb)[TN], int kb) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
      fa[a] = *(const i4v*)(Ab + ((int64_t)a * nkb + kb) * 1024);
#pragma unroll
    for (int b = 0; b < TN; ++b)
      fb[b] = *(const i4v*)(Bb + ((int64_t)b * nkb + kb) * 1024);
  };
  auto mma = [&](const i4v(&fa)[TM], const i4v(&fb)[TN]) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[a], fb[b],
                                                          acc[a][b], 0, 0, 0);
  };
  if (nk > 0) {
    load(fa0, fb0, 0);
    int kb = 0;
    for (; kb + 1 < nk; kb += 2) {
      load(fa1, fb1, kb + 1);
      mma(fa0, fb0);
      if (kb + 2 < nk) load(fa0, fb0, kb + 2);
      mma(fa1, fb1);
    }
    if (kb < nk) mma(fa0, fb0);
  }

  // |acc| <= K 127^2 < 2^31: exact in double
  const double m = (double)cc.m[mi], minv = cc.minv_d[mi];
  int8_t* R = g.R + (int64_t)z * nr * ldr;
#pragma unroll
  for (int b = 0; b < TN; ++b) {
    const int n = (brb + b) * 32 + (lane & 31);
#pragma unroll
    for (int a = 0; a < TM; ++a)
      store_residue_tile(acc[a][b], m, minv,
                         R + ((int64_t)(arb + a - rb0) * nr + n) * 32,
                         n < g.N, lane);
  }
}

// Same product through LDS: NST stages of KPB k blocks of the A and B
// blocks of the workgroup tile, filled by direct global -> LDS loads (each
// wave-wide global_load_lds_dwordx4 copies one 1 KB block; lane l fetches
// chunk sigma(l) so the fragment reads are bank-conflict-free), one
// barrier per stage.  The deep prefetch costs no registers, which matters
// here: a k block carries only TM x TN products per wave.
//   PF = false (KPB = 1): stage kb + NST - 1 is requested, then the
//     fragments of stage kb are read and multiplied.
//   PF = true: the fragments of stage s + 1 are read from LDS (into a
//     second register set) while the products of stage s are issued, so
//     the LDS latency hides behind the matrix pipe; stage s + 1 must have
//     landed by then, so NST - 2 stages stay in flight.  KPB = 2 halves the
//     barriers per k block (the k block count must then be even).
template <int TM, int TN, int WM, int WN, int NST, bool PF, int KPB>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm_i8_res_lds(
    I8GemmArgs g, CrtConsts cc, int ntm, int ntn, int rb0, int nkb,
    int64_t ldr, int nr) {
  constexpr int NW = WM * WN;
  constexpr int BMB = TM * WM, BNB = TN * WN;
  constexpr int NBLK = BMB + BNB;
  constexpr int NBLKP = (NBLK + NW - 1) / NW * NW;
  constexpr int DPW = NBLKP / NW;
  constexpr int SB = NBLKP * KPB;  // blocks per stage
  static_assert(NST >= 2 && NST <= 6 && (NST - 2) * DPW * KPB < 64,
                "stage count");
  static_assert(!PF || NST >= 3, "fragment prefetch needs 3 stages");
  static_assert(KPB == 1 || (PF && KPB == 2), "k blocks per stage");
  __shared__ __attribute__((aligned(1024))) int8_t lds[NST * SB * 1024];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave % WM, wn = wave / WM;
  int za, mi, tm, tn;
  crt_tile(g, ntm, ntn, &za, &mi, &tm, &tn);
  const int z = za * g.S + mi;
  const int arb0 = rb0 + tm * BMB, brb0 = tn * BNB;
  // k blocks [k0, k0 + nk) of the tile (all nkb without A ranges); the
  // operand pointers start at block k0, rows stay nkb blocks long
  int k0, nk;
  crt_tile_krange(g, arb0, BMB, nkb, 2, &k0, &nk);
  const int src_off = 16 * i8_sigma(lane);
  const int8_t* Az = g.A + za * g.a_z + mi * g.a_plane + (int64_t)k0 * 1024;
  const int8_t* Bz = g.B + mi * g.b_plane + (int64_t)k0 * 1024;
  const int nst = nk / KPB;  // stages
  // slot j * NBLKP + q of a stage: k block j of the stage, A blocks then B
  // blocks
  auto issue = [&](int stage, int st) {
    int8_t* dst = lds + stage * (SB * 1024);
#pragma unroll
    for (int j = 0; j < KPB; ++j)
#pragma unroll
      for (int t = 0; t < DPW; ++t) {
        const int kb = st * KPB + j;
        const int q = wave + NW * t;
        const int qq = q < NBLK ? q : 0;  // padding slots repeat block 0
        const int8_t* src =
            qq < BMB ? Az + ((int64_t)(arb0 + qq) * nkb + kb) * 1024
                     : Bz + ((int64_t)(brb0 + qq - BMB) * nkb + kb) * 1024;
        __builtin_amdgcn_global_load_lds(
            (const void*)(src + src_off),
            (__attribute__((address_space(3))) void*)(dst +
                                                      (j * NBLKP + q) * 1024),
            16, 0, 0);
      }
  };
  // request idx = j * DPW + t of those issue() makes
  auto issue_one = [&](int stage, int st, int idx) {
    const int j = idx / DPW, t = idx - j * DPW;
    int8_t* dst = lds + stage * (SB * 1024);
    const int kb = st * KPB + j;
    const int q = wave + NW * t;
    const int qq = q < NBLK ? q : 0;
    const int8_t* src =
        qq < BMB ? Az + ((int64_t)(arb0 + qq) * nkb + kb) * 1024
                 : Bz + ((int64_t)(brb0 + qq - BMB) * nkb + kb) * 1024;
    __builtin_amdgcn_global_load_lds(
        (const void*)(src + src_off),
        (__attribute__((address_space(3))) void*)(dst +
                                                  (j * NBLKP + q) * 1024),
        16, 0, 0);
  };
  i16v acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0;
  const int rd_off = 16 * i8_sigma(2 * (lane & 31) + (lane >> 5));
  auto read = [&](i4v(&fa)[TM], i4v(&fb)[TN], int stage, int j) {
    const int8_t* base = lds + (stage * SB + j * NBLKP) * 1024 + rd_off;
#pragma unroll
    for (int a = 0; a < TM; ++a)
      fa[a] = *(const i4v*)(base + (TM * wm + a) * 1024);
#pragma unroll
    for (int b = 0; b < TN; ++b)
      fb[b] = *(const i4v*)(base + (BMB + TN * wn + b) * 1024);
  };
  // products of A fragments [a0, a1)
  auto mma = [&](const i4v(&fa)[TM], const i4v(&fb)[TN], int a0 = 0,
                 int a1 = TM) {
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
      for (int b = 0; b < TN; ++b)
        if (a >= a0 && a < a1)
          acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
              fa[a], fb[b], acc[a][b], 0, 0, 0);
  };
  // the epilogue's scalars, loaded (and waited for) before the loop, so
  // that no scalar load is in flight inside it (scalar and LDS returns
  // share one counter)
  int8_t* R = g.R + (int64_t)z * nr * ldr;
  double m = (double)cc.m[mi], minv = cc.minv_d[mi];
  asm volatile("" : "+v"(R), "+v"(m), "+v"(minv));
#pragma unroll
  for (int st = 0; st < NST - 1; ++st)
    if (st < nst) issue(st, st);
  if constexpr (!PF) {
    int stage = 0;
    int kb = 0;
    // k blocks that request a stage, in a loop of their own: the fragment
    // reads first, then the requests spread over the first products, one
    // after each (requests issued between the barrier and the reads kept
    // the matrix pipe idle longer: the half-ls GEMM ran 7% slower)
    for (; kb + NST - 1 < nk; ++kb) {
      wait_stages<DPW, NST - 2>(NST - 2);
      __builtin_amdgcn_s_barrier();
      i4v fa[TM], fb[TN];
      read(fa, fb, stage, 0);
      const int slot = stage == 0 ? NST - 1 : stage - 1;
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) {
          acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
              fa[a], fb[b], acc[a][b], 0, 0, 0);
          if (a * TN + b < DPW) {
            __builtin_amdgcn_sched_barrier(0);
            issue_one(slot, kb + NST - 1, a * TN + b);
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      stage = stage == NST - 1 ? 0 : stage + 1;
    }
    for (; kb < nk; ++kb) {
      wait_stages<DPW, NST - 2>(min(NST - 2, nk - 1 - kb));
      __builtin_amdgcn_s_barrier();
      if (kb + NST - 1 < nk)
        issue(stage == 0 ? NST - 1 : stage - 1, kb + NST - 1);
      i4v fa[TM], fb[TN];
      read(fa, fb, stage, 0);
      mma(fa, fb);
      stage = stage == NST - 1 ? 0 : stage + 1;
    }
  } else if (nst > 0) {
    i4v fa0[KPB][TM], fb0[KPB][TN], fa1[KPB][TM], fb1[KPB][TN];
    wait_stages<DPW * KPB, NST - 2>(min(NST - 2, nst - 1));
    __builtin_amdgcn_s_barrier();
#pragma unroll
    for (int j = 0; j < KPB; ++j) read(fa0[j], fb0[j], 0, j);
    // iteration s: stage s + 1 landed -> barrier -> request stage
    // s + NST - 1 into the buffer of stage s - 1 (read before the previous
    // barrier and consumed by products that precede this one) -> products
    // of stage s, with the reads of stage s + 1 among them.
    // With LDS DMA in flight the LDS counter is not in order, so the wait
    // before a product whose fragments came from LDS drains every LDS read:
    // the reads of k block j of stage s + 1 are therefore issued after the
    // first half of the products of k block j of stage s (whose wait then
    // finds the reads made half a k block earlier complete) and before the
    // second half.  The steady-state loop reads unconditionally (a read on
    // one path only would also force a full drain).
    int nxt = 1 % NST;  // buffer of stage s + 1
    auto step = [&](const i4v(&ca)[KPB][TM], const i4v(&cb)[KPB][TN],
                    i4v(&na)[KPB][TM], i4v(&nb)[KPB][TN], int st) {
      wait_stages<DPW * KPB, NST - 3>(min(NST - 3, nst - 2 - st));
      __builtin_amdgcn_s_barrier();
      if (st + NST - 1 < nst)
        issue(nxt == 0 ? NST - 2 : (nxt == 1 ? NST - 1 : nxt - 2),
              st + NST - 1);
#pragma unroll
      for (int j = 0; j < KPB; ++j) {
        mma(ca[j], cb[j], 0, (TM + 1) / 2);
        __builtin_amdgcn_sched_barrier(0);
        read(na[j], nb[j], nxt, j);
        __builtin_amdgcn_sched_barrier(0);
        mma(ca[j], cb[j], (TM + 1) / 2, TM);
      }
      nxt = nxt == NST - 1 ? 0 : nxt + 1;
    };
    // The steps that request a stage (all but the last NST - 1) run in a
    // loop of their own in which every step requests unconditionally, the
    // requests spread over the first products of the step, one after each,
    // so that a wave held while its requests are accepted still has
    // products in its matrix pipe (the requests issued back to back after
    // the barrier leave the pipe idle meanwhile: the projection runs 4%
    // slower that way).  A request under a branch among the products would
    // make the two paths keep the accumulators in different registers,
    // hence the separate loop.
    auto step_ilv = [&](const i4v(&ca)[KPB][TM], const i4v(&cb)[KPB][TN],
                        i4v(&na)[KPB][TM], i4v(&nb)[KPB][TN], int st) {
      constexpr int HALF = (TM + 1) / 2 * TN;  // first-half products
      static_assert(DPW * KPB <= HALF, "one request per product");
      wait_stages<DPW * KPB, NST - 3>(NST - 3);
      __builtin_amdgcn_s_barrier();
      const int slot = nxt == 0 ? NST - 2 : (nxt == 1 ? NST - 1 : nxt - 2);
#pragma unroll
      for (int j = 0; j < KPB; ++j) {
        if (j == 0) {
          // request idx after product idx of the first k block
#pragma unroll
          for (int a = 0; a < (TM + 1) / 2; ++a)
#pragma unroll
            for (int b = 0; b < TN; ++b) {
              acc[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(
                  ca[0][a], cb[0][b], acc[a][b], 0, 0, 0);
              if (a * TN + b < DPW * KPB) {
                __builtin_amdgcn_sched_barrier(0);
                issue_one(slot, st + NST - 1, a * TN + b);
                __builtin_amdgcn_sched_barrier(0);
              }
            }
        } else {
          mma(ca[j], cb[j], 0, (TM + 1) / 2);
        }
        __builtin_amdgcn_sched_barrier(0);
        read(na[j], nb[j], nxt, j);
        __builtin_amdgcn_sched_barrier(0);
        mma(ca[j], cb[j], (TM + 1) / 2, TM);
      }
      nxt = nxt == NST - 1 ? 0 : nxt + 1;
    };
    int st = 0;
    for (; st + 2 < nst && st + NST < nst; st += 2) {
      step_ilv(fa0, fb0, fa1, fb1, st);
      step_ilv(fa1, fb1, fa0, fb0, st + 1);
    }
    for (; st + 2 < nst; st += 2) {
      step(fa0, fb0, fa1, fb1, st);
      step(fa1, fb1, fa0, fb0, st + 1);
    }
    // one or two stages left
    if (st + 1 < nst) {
      step(fa0, fb0, fa1, fb1, st);
#pragma unroll
      for (int j = 0; j < KPB; ++j) mma(fa1[j], fb1[j]);
    } else {
#pragma unroll
      for (int j = 0; j < KPB; ++j) mma(fa0[j], fb0[j]);
    }
  }

  const int arb = arb0 + TM * wm, brb = brb0 + TN * wn;
#pragma unroll
  for (int b = 0; b < TN; ++b) {
    const int n = (brb + b) * 32 + (lane & 31);
#pragma unroll
    for (int a = 0; a < TM; ++a)
      store_residue_tile(acc[a][b], m, minv,
                         R + ((int64_t)(arb + a - rb0) * nr + n) * 32,
                         n < g.N, lane);
  }
}

// The integer product of the operands from the nm residues of four outputs
// (byte q of word w[i] = symmetric residue of output q modulo m_i), in
// FP64.
//   nm <= 12, explicit form (CrtConsts): q = rint(sum r_i ef_i) (its error is
//   below 2^-39 while |X| / M < 1/2 - 3e-7, crt_beta), then the high parts
//   sum_i r_i ech_i - q emh exactly and the low parts with rounding; X is
//   within 2^12 + 2 ulp of the exact product (the operand rounding puts
//   >= 2^44 into every output).
//   nm = 14, by pairs: each pair of moduli joined exactly in FP32, Garner's
//   digits over the pair moduli in FP64 (exact integers below 2^33), the
//   mixed-radix number by Horner in FP64 (within 4 ulp).
template <int NM>
__host__ __device__ inline void crt_recon4(const uint32_t (&w)[NM],
                                           const CrtConsts& cc,
                                           double (&X)[4]) {
  if constexpr (NM <= 12) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double r[NM];
#pragma unroll
      for (int i = 0; i < NM; ++i)
        r[i] = (double)(float)(int8_t)(w[i] >> (8 * q));
      double e = 0.0;
#pragma unroll
      for (int i = 0; i < NM; ++i) e = fma(r[i], cc.ef[i], e);
      const double qt = rint(e);
      double hi = -qt * cc.emh, lo = -qt * cc.eml;
#pragma unroll
      for (int i = 0; i < NM; ++i) {
        hi = fma(r[i], cc.ech[i], hi);
        lo = fma(r[i], cc.ecl[i], lo);
      }
      X[q] = hi + lo;
    }
    return;
  }
  constexpr int NP = (NM + 1) / 2;
  double v[NP][4];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const int ia = 2 * k, ib = 2 * k + 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float ra = (float)(int8_t)(w[ia] >> (8 * q));
      float r = ra;
      if (ib < NM) {
        const float rb = (float)(int8_t)(w[ib] >> (8 * q));
        const float d = sym_mod_f((rb - ra) * cc.pinv[k], (float)cc.m[ib],
                                  cc.minv[ib]);
        r = fmaf((float)cc.m[ia], d, ra);
      }
      if (k == 0) {
        v[0][q] = (double)r;
      } else {
        double t = (double)r * cc.pgi[k];
#pragma unroll
        for (int j = 0; j < k; ++j) t = fma(v[j][q], cc.pnw[k][j], t);
        v[k][q] = sym_mod_d(t, cc.P[k], cc.Pinv[k]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double x = v[NP - 1][q];
#pragma unroll
    for (int k = NP - 2; k >= 0; --k) x = fma(x, cc.P[k], v[k][q]);
    X[q] = x;
  }
}

// Reconstruction of every output of a CRT-form GEMM into FP64, scaled by
// the operand exponents.  Four consecutive outputs per thread (one 32-bit
// load per residue plane; residue rows are 4-aligned).
template <int NM>
__global__ __launch_bounds__(256) void k_crt_combine(I8GemmArgs g,
                                                     CrtConsts cc, int rb0,
                                                     int64_t ldr, int nr) {
  // workgroup: one residue row block x 32 columns; thread: column
  // t / 8, rows 4 (t % 8) .. + 3, so a wave reads 256 contiguous bytes of
  // each residue plane
  const int t = threadIdx.x;
  const int n = blockIdx.y * 32 + (t >> 3);
  const int za = blockIdx.z;
  const int r0 = blockIdx.x * 32 + 4 * (t & 7);  // residue row, 4-aligned
  const int rlo = g.row0 - rb0 * 32;  // residue row of output 0
  if (n >= g.N || r0 >= rlo + g.M || r0 + 4 <= rlo) return;
  const uint32_t* R = (const uint32_t*)(g.R + (int64_t)za * NM * nr * ldr +
                                        ((int64_t)blockIdx.x * nr + n) * 32 +
                                        4 * (t & 7));
  const int64_t pst = (int64_t)nr * ldr / 4;
  uint32_t w[NM];
#pragma unroll
  for (int i = 0; i < NM; ++i) w[i] = R[i * pst];
  double X[4];
  crt_recon4<NM>(w, cc, X);
  const int eb = g.eB[n] - 2 * (g.beta - 1);
  double* C = g.C + za * g.c_z + (int64_t)n * g.ldc;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int m = r0 + q - rlo;
    if (m < 0 || m >= g.M) continue;
    C[m] = ldexp(X[q], g.eA[za * g.ea_z + g.row0 + m] + eb);
  }
}

// Fused reconstruction and K-block gather of the half line strength.  One
// workgroup per (32-row block of v_F, vector): phase 1 reconstructs the
// block's Y_alpha[col][32 rows] for the three components and every X
// column into LDS exactly as k_crt_combine does (a wave reads 256
// contiguous bytes of each residue plane); phase 2 forms every final
// (K_F, tau_F, v_F) of the block from LDS with the rotational factors and
// the term order of k_gather_h, so H is bit for bit that of the two-kernel
// path, without the FP64 Y round trip through HBM.  The rotational factors
// of every (final column, K_I, tau_I) are tabulated in LDS once per
// workgroup (their loads overlap phase 1), and phase 2 takes four outputs
// per thread at a time, their slot lookups in flight together.
constexpr int kCrtGatherMaxColF = 64;  // final columns 2 (J_F + 1), J_F < 32
constexpr int kCrtGatherVecs = 4;      // vectors per workgroup (grid stride)

template <int NM>
__global__ __launch_bounds__(256) void k_crt_gather_h(CrtGatherArgs a,
                                                      CrtConsts cc,
                                                      int64_t ldr, int nr) {
  extern __shared__ double ys[];  // [3][ncolI][32]
  // factor and component of term (K_I = K_F - 1 + q, tau_I) of final
  // column colF at [colF * 6 + 2 q + tau_I] (factor 0: no term)
  __shared__ double cf[kCrtGatherMaxColF * 6];
  __shared__ int8_t cal[kCrtGatherMaxColF * 6];
  const I8GemmArgs& g = a.g;
  const int t = threadIdx.x;
  const int rb = blockIdx.x;
  const int ncolI = a.ncolI;
  const int ncolF = 2 * (a.jF + 1);
  const int64_t pst = (int64_t)nr * ldr / 4;
  const int64_t zst = (int64_t)NM * nr * ldr;  // between components
  const int n1 = a.jmax + 1;
  for (int e = t; e < ncolF * 6; e += 256) {
    const int colF = e / 6, q = (e % 6) >> 1, tI = e & 1;
    const int K = colF >> 1, tF = colF & 1;
    const int KI = K - 1 + q;
    double coef = 0.0;
    int al = 0;
    if (KI >= 0 && KI <= a.jI) {
      const int dk = K - KI;
      const double f3j = a.threej[a.jI + KI * n1 + (a.jF - a.jI + 1) * n1 * n1 +
                                  (dk + 1) * n1 * n1 * 3];
      const double sq = (dk != 0 && KI * K != 0) ? kSqrt2Ref : 1.0;
      int m;
      if (dk == 0) {
        al = 2;
        m = tF - tI;
      } else if (tF != tI) {
        al = 0;
        m = (tF - tI) * dk;
      } else {
        al = 1;
        m = -1;
      }
      coef = f3j * sq * (double)m;
    }
    cf[e] = coef;
    cal[e] = (int8_t)al;
  }
  const int nitem1 = 3 * ncolI * 8;            // (component, column, quad)
  const int nitem2 = ncolF * 32;               // (final column, row)
  for (int vec = blockIdx.y; vec < a.nvec; vec += gridDim.y) {
    const int64_t colv = (int64_t)vec * ncolI;
    // two items per thread at a time: every residue word and exponent of
    // both is requested before either is reconstructed
#pragma unroll 1
    for (int i0 = t; i0 < nitem1; i0 += 2 * 256) {
      uint32_t w[2][NM];
      int ex[2][4];
      int off[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = min(i0 + h * 256, nitem1 - 1);
        const int al = i / (ncolI * 8);
        const int rem = i - al * ncolI * 8;
        const int col = rem >> 3, quad = rem & 7;
        const int64_t n = colv + col;
        const uint32_t* R =
            (const uint32_t*)(g.R + al * zst + ((int64_t)rb * nr + n) * 32 +
                              4 * quad);
#pragma unroll
        for (int k = 0; k < NM; ++k) w[h][k] = R[k * pst];
        const int eb = g.eB[n] - 2 * (g.beta - 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int v = min(rb * 32 + 4 * quad + q, a.nc - 1);
          ex[h][q] = g.eA[al * g.ea_z + v] + eb;
        }
        off[h] = (al * ncolI + col) * 32 + 4 * quad;
      }
      // keep both items' requests ahead of the reconstructions
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (i0 + h * 256 >= nitem1) break;
        double X[4];
        crt_recon4<NM>(w[h], cc, X);
        double* yo = ys + off[h];
#pragma unroll
        for (int q = 0; q < 4; ++q) yo[q] = ldexp(X[q], ex[h][q]);
      }
    }
    __syncthreads();
    double* Hv = a.H + (int64_t)vec * a.hstride;
#pragma unroll 1
    for (int j0 = t; j0 < nitem2; j0 += 4 * 256) {
      int pr[4];
      double val[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int j = j0 + u * 256;
        const int colF = j >> 5, r = j & 31;
        const int v = rb * 32 + r;
        pr[u] = -1;
        val[u] = 0.0;
        if (j >= nitem2 || v >= a.nc) continue;
        pr[u] = a.slot_one[colF * a.nc + v];
        const int K = colF >> 1, tF = colF & 1;
        double acc = 0.0;
#pragma unroll
        for (int e = 0; e < 6; ++e) {
          const double coef = cf[colF * 6 + e];
          if (coef == 0.0) continue;
          const int KI = K - 1 + (e >> 1);
          double y = 0.0;  // as k_gather_h sums its (single) split
          y += ys[(cal[colF * 6 + e] * ncolI + 2 * KI + (e & 1)) * 32 + r];
          acc = fma(coef, y, acc);
        }
        const bool neg = ((K % 3) * tF) & 1;
        val[u] = neg ? -acc : acc;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (pr[u] >= 0) {
          Hv[pr[u]] = val[u];
        } else if (pr[u] == -2) {  // several primitives share the slot
          const int j = j0 + u * 256;
          const int sl = (j >> 5) * a.nc + rb * 32 + (j & 31);
          for (int p = a.slot_ptr[sl]; p < a.slot_ptr[sl + 1]; ++p)
            Hv[a.slot_prim[p]] = val[u];
        }
      }
    }
    __syncthreads();  // before the next vector overwrites ys
  }
}

namespace {

struct CrtLaunch {
  int ntm, ntn, rb0, nkb, nr;
  int64_t grid, ldr;
};

CrtLaunch crt_launch(const I8GemmArgs& g) {
  const I8Cfg& cf = kCrtCfg[g.cfg];
  const int bm = 32 * cf.tm * cf.wm, bn = 32 * cf.tn * cf.wn;
  CrtLaunch L;
  L.rb0 = g.row0 / 32;
  const int rows = g.row0 + g.M - L.rb0 * 32;
  L.ntm = (rows + bm - 1) / bm;
  L.ntn = (g.N + bn - 1) / bn;
  L.nkb = (int)(g.Kp / 32);
  L.grid = (int64_t)L.ntm * L.ntn * g.nz * g.S;
  L.ldr = (int64_t)L.ntm * bm;  // residue rows per column
  L.nr = L.ntn * bn;            // residue columns per matrix
  return L;
}

}  // namespace

// One workgroup per (matrix, row block): each thread ORs 16-byte chunks of
// the row block's k blocks over every plane; the first and last k block
// with a nonzero chunk are kept by LDS atomics.
__global__ __launch_bounds__(256) void k_block_kranges(
    const int8_t* __restrict__ planes, int64_t plane, int nplanes,
    int64_t mat_stride, int nrb, int nkb, int2* __restrict__ out) {
  __shared__ int lo_s, hi_s;
  if (threadIdx.x == 0) {
    lo_s = nkb;
    hi_s = 0;
  }
  __syncthreads();
  const int rb = blockIdx.x, mat = blockIdx.y;
  const int8_t* base =
      planes + mat * mat_stride + (int64_t)rb * nkb * 1024;
  int lo = nkb, hi = 0;
  for (int idx = threadIdx.x; idx < nkb * 64; idx += 256) {
    const int kb = idx >> 6;
    const int8_t* p = base + (int64_t)idx * 16;
    uint4 o = make_uint4(0, 0, 0, 0);
    for (int q = 0; q < nplanes; ++q) {
      const uint4 v = *(const uint4*)(p + q * plane);
      o.x |= v.x;
      o.y |= v.y;
      o.z |= v.z;
      o.w |= v.w;
    }
    if (o.x | o.y | o.z | o.w) {
      lo = min(lo, kb);
      hi = max(hi, kb + 1);
    }
  }
  if (lo < hi) {
    atomicMin(&lo_s, lo);
    atomicMax(&hi_s, hi);
  }
  __syncthreads();
  if (threadIdx.x == 0)
    out[(int64_t)mat * nrb + rb] =
        lo_s < hi_s ? make_int2(lo_s, hi_s) : make_int2(0, 0);
}

void launch_crt_gather_h(const CrtGatherArgs& a, hipStream_t s) {
  const I8GemmArgs& g = a.g;
  if (a.nvec <= 0 || a.nc <= 0) return;
  GAIN_CHECK_ARG(g.nz == 3 && g.row0 == 0 && g.M == a.nc &&
                     g.N == a.nvec * a.ncolI && g.R &&
                     a.ncolI == 2 * (a.jI + 1),
                 "CRT gather: product shape (nz %d, row0 %d, M %d, N %d) "
                 "does not match %d vectors of %d columns", g.nz, g.row0,
                 g.M, g.N, a.nvec, a.ncolI);
  GAIN_CHECK_ARG(a.H && a.slot_ptr && a.slot_prim && a.slot_one &&
                     a.threej && g.eA && g.eB && a.hstride > 0 &&
                     2 * (a.jF + 1) <= kCrtGatherMaxColF,
                 "CRT gather: missing output, final slots, 3j table or "
                 "exponents");
  const CrtLaunch L = crt_launch(g);
  const CrtConsts& cc = crt_consts(g.S);
  // LDS of one (row block, vector): Y of 3 components x every X column
  const size_t lds = (size_t)3 * a.ncolI * 32 * sizeof(double);
  GAIN_CHECK_ARG(lds <= 64 * 1024, "CRT gather: %d X columns exceed the LDS "
                 "tile", a.ncolI);
  // several vectors per workgroup share its coefficient table
  const dim3 grid((unsigned)((a.nc + 31) / 32),
                  (unsigned)std::min((a.nvec + kCrtGatherVecs - 1) /
                                         kCrtGatherVecs,
                                     65535));
  switch (g.S) {
#define GAIN_CRT_GATHER(NN)                                                  \
  case NN:                                                                   \
    hipLaunchKernelGGL(k_crt_gather_h<NN>, grid, dim3(256), lds, s, a, cc,   \
                       L.ldr, L.nr);                                         \
    break;
    GAIN_CRT_GATHER(11)
    GAIN_CRT_GATHER(12)
    GAIN_CRT_GATHER(14)
#undef GAIN_CRT_GATHER
    default:
      fail(GAIN_ERR_ARG, "CRT gather: %d moduli not supported", g.S);
  }
  check_launch("crt_gather_h");
}

void i8_block_kranges(const int8_t* planes, int64_t plane, int nplanes,
                      int64_t mat_stride, int nmat, int nrb, int nkb,
                      int2* out, hipStream_t s) {
  if (nmat <= 0 || nrb <= 0) return;
  GAIN_CHECK_ARG(nplanes >= 1 && nkb >= 1 && nmat <= 65535,
                 "k-block ranges: %d planes, %d k blocks, %d matrices",
                 nplanes, nkb, nmat);
  hipLaunchKernelGGL(k_block_kranges, dim3(nrb, nmat), dim3(256), 0, s,
                     planes, plane, nplanes, mat_stride, nrb, nkb, out);
  check_launch("block_kranges");
}

double crt_kblock_fraction(const I8GemmArgs& g, const int2* kr) {
  if (!g.a_kr || !kr) return 1.0;
  const CrtLaunch L = crt_launch(g);
  const int bmb = kCrtCfg[g.cfg].tm * kCrtCfg[g.cfg].wm;
  int64_t done = 0, dense = 0;
  for (int tm = 0; tm < L.ntm; ++tm) {
    // as crt_tile_krange: one range for the nz matrices of a tile row
    int lo = L.nkb, hi = 0;
    for (int z = 0; z < g.nz; ++z)
      for (int q = 0; q < bmb; ++q) {
        const int2 r = kr[z * g.akr_z + L.rb0 + tm * bmb + q];
        if (r.x < r.y) {
          lo = std::min(lo, r.x);
          hi = std::max(hi, r.y);
        }
      }
    if (lo < hi)
      done += g.nz * (std::min(L.nkb, (hi + 1) / 2 * 2) - lo / 2 * 2);
    dense += g.nz * L.nkb;
  }
  return dense ? (double)done / (double)dense : 1.0;
}

int64_t gemm_crt_res_bytes(const I8GemmArgs& g) {
  const CrtLaunch L = crt_launch(g);
  return (int64_t)g.nz * g.S * L.nr * L.ldr;
}

int64_t gemm_crt_grid_items(const I8GemmArgs& g) {
  return crt_launch(g).grid * 64 * kCrtCfg[g.cfg].wm * kCrtCfg[g.cfg].wn;
}

void gemm_crt_i8(const I8GemmArgs& g, hipStream_t s) {
  gemm_crt_res(g, s);
  gemm_crt_combine(g, s);
}

void gemm_crt_res(const I8GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  GAIN_CHECK_ARG(g.cfg >= 0 && g.cfg < kCrtConfigs, "CRT gemm: config %d",
                 g.cfg);
  GAIN_CHECK_ARG(g.Kp > 0 && g.Kp % 64 == 0 && g.ksplit == 1 && g.nz >= 1 &&
                     g.row0 >= 0 && g.R,
                 "CRT gemm: Kp %lld, split %d, residue buffer %p",
                 (long long)g.Kp, g.ksplit, (const void*)g.R);
  // exact int32 sums per modulus, and an integer width the moduli cover
  GAIN_CHECK_ARG(g.Kp * 127 * 127 < (1ll << 31),
                 "CRT gemm: K = %lld beyond the exact int32 range",
                 (long long)g.Kp);
  GAIN_CHECK_ARG(g.beta >= 9 && g.beta <= crt_beta(g.S, g.Kp),
                 "CRT gemm: beta %d too wide for K = %lld and %d moduli",
                 g.beta, (long long)g.Kp, g.S);
  const CrtLaunch L = crt_launch(g);
  GAIN_CHECK_ARG(L.grid < (1ll << 31), "CRT gemm: grid too large");
  // every tile's A row blocks must have a range entry
  GAIN_CHECK_ARG(!g.a_kr ||
                     (g.akr_z >= g.akr_n &&
                      L.rb0 + L.ntm * kCrtCfg[g.cfg].tm * kCrtCfg[g.cfg].wm <=
                          g.akr_n),
                 "CRT gemm: k-block ranges cover %d row blocks, tiles reach "
                 "%d", g.akr_n,
                 L.rb0 + L.ntm * kCrtCfg[g.cfg].tm * kCrtCfg[g.cfg].wm);
  const CrtConsts& cc = crt_consts(g.S);
  const dim3 grid((unsigned)L.grid),
      block(64 * kCrtCfg[g.cfg].wm * kCrtCfg[g.cfg].wn);
  switch (g.cfg) {
#define GAIN_CRT_CASE(CC)                                                    \
  case CC:                                                                   \
    hipLaunchKernelGGL((k_gemm_i8_res<kCrtCfg[CC].tm, kCrtCfg[CC].tn,        \
                                      kCrtCfg[CC].wm, kCrtCfg[CC].wn>),      \
                       grid, block, 0, s, g, cc, L.ntm, L.ntn, L.rb0, L.nkb, \
                       L.ldr, L.nr);                                         \
    break;
    GAIN_CRT_CASE(0)
    GAIN_CRT_CASE(1)
    GAIN_CRT_CASE(2)
    GAIN_CRT_CASE(3)
    GAIN_CRT_CASE(4)
#undef GAIN_CRT_CASE
#define GAIN_CRT_LDS(CC)                                                     \
  case CC:                                                                   \
    hipLaunchKernelGGL((k_gemm_i8_res_lds<                                  \
                           kCrtCfg[CC].tm, kCrtCfg[CC].tn, kCrtCfg[CC].wm,   \
                           kCrtCfg[CC].wn, crt_nst(kCrtCfg[CC].lds),         \
                           crt_pf(kCrtCfg[CC].lds),                          \
                           crt_kpb(kCrtCfg[CC].lds)>),                       \
                       grid, block, 0, s, g, cc, L.ntm, L.ntn, L.rb0, L.nkb, \
                       L.ldr, L.nr);                                         \
    break;
    GAIN_CRT_LDS(5)
    GAIN_CRT_LDS(6)
    GAIN_CRT_LDS(7)
    GAIN_CRT_LDS(8)
    GAIN_CRT_LDS(9)
    GAIN_CRT_LDS(10)
    GAIN_CRT_LDS(11)
    GAIN_CRT_LDS(12)
    GAIN_CRT_LDS(13)
#undef GAIN_CRT_LDS
  }
  check_launch("gemm_i8_res");
}

// The same reconstruction on the host (CPU tests of the constants and the
// arithmetic): words [n4][nm], out [n4][4].
template <int NM>
static void crt_recon_host(const uint32_t* words, int64_t n4, double* out) {
  const CrtConsts& cc = crt_consts(NM);
  for (int64_t g = 0; g < n4; ++g) {
    uint32_t w[NM];
    for (int i = 0; i < NM; ++i) w[i] = words[g * NM + i];
    double X[4];
    crt_recon4<NM>(w, cc, X);
    for (int q = 0; q < 4; ++q) out[g * 4 + q] = X[q];
  }
}

void crt_reconstruct_host(int nm, const uint32_t* words, int64_t n4,
                          double* out) {
  switch (nm) {
    case 11: crt_recon_host<11>(words, n4, out); break;
    case 12: crt_recon_host<12>(words, n4, out); break;
    case 14: crt_recon_host<14>(words, n4, out); break;
    default: fail(GAIN_ERR_ARG, "CRT: %d moduli not supported", nm);
  }
}

int crt_modulus(int nm, int i) {
  GAIN_CHECK_ARG(i >= 0 && i < nm && (nm == 11 || nm == 12 || nm == 14),
                 "CRT modulus %d of %d", i, nm);
  return crt_consts(nm).m[i];
}

void gemm_crt_combine(const I8GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  GAIN_CHECK_ARG(g.C && g.R && g.eA && g.eB,
                 "CRT combine: output %p, residues %p, exponents %p / %p",
                 (const void*)g.C, (const void*)g.R, (const void*)g.eA,
                 (const void*)g.eB);
  const CrtLaunch L = crt_launch(g);
  const CrtConsts& cc = crt_consts(g.S);
  // residue rows [row0 - 32 rb0, row0 - 32 rb0 + M): one workgroup per
  // row block and 32 columns
  const int rspan = g.row0 - L.rb0 * 32 + g.M;
  const dim3 cgrid((unsigned)((rspan + 31) / 32), (unsigned)((g.N + 31) / 32),
                   (unsigned)g.nz);
  switch (g.S) {
#define GAIN_CRT_COMB(NN)                                                    \
  case NN:                                                                   \
    hipLaunchKernelGGL(k_crt_combine<NN>, cgrid, dim3(256), 0, s, g, cc,     \
                       L.rb0, L.ldr, L.nr);                                  \
    break;
    GAIN_CRT_COMB(11)
    GAIN_CRT_COMB(12)
    GAIN_CRT_COMB(14)
#undef GAIN_CRT_COMB
    default:
      fail(GAIN_ERR_ARG, "CRT gemm: %d moduli not supported", g.S);
  }
  check_launch("crt_combine");
}

}  // namespace gain
