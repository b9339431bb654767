// Split-BF16 slicing and sliced GEMM (v_mfma_f32_16x16x32_bf16).  The
// scheme and its exactness bounds are described in ozaki.hpp.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "ozaki.hpp"

namespace gain {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void k_slice_rows(
    const double* __restrict__ src, int64_t ld, int rows, int K, int64_t Kp,
    uint16_t* __restrict__ planes, int* __restrict__ expo) {
  constexpr int CH = (S <= 3 ? 256 : 128);
  constexpr int WPC = CH / 64;  // waves per chunk
  __shared__ double red[4];
  const int r = blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6;
  const double* x = src + (int64_t)r * ld;
  const int64_t pstride = (int64_t)rows * Kp;
  const int64_t nch = Kp / CH;
  uint16_t* out = planes + (int64_t)r * Kp;
  for (int64_t k0 = 0; k0 < Kp; k0 += 256) {
    const int64_t k = k0 + tid;
    double v = k < K ? x[k] : 0.0;
    double mx = fabs(v);
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
    __syncthreads();
    if ((tid & 63) == 0) red[wave] = mx;
    __syncthreads();
    const int w0 = wave / WPC * WPC;
    mx = red[w0];
#pragma unroll
    for (int w = 1; w < WPC; ++w) mx = fmax(mx, red[w0 + w]);
    int e = 0;
    if (mx > 0.0) frexp(mx, &e);  // mx = f 2^e, f in [0.5, 1): |x| 2^-e < 1
    if (tid % CH == 0) expo[(int64_t)r * nch + k / CH] = e;
    v = ldexp(v, -e);
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const double n = rint(ldexp(v, 8 * (i + 1)));
      const double sl = ldexp(n, -8 * (i + 1));
      v -= sl;  // exact: sl is v rounded to the slice grid
      const float f = (float)sl;  // <= 8 significant bits: exact
      out[i * pstride + k] = (uint16_t)(__float_as_uint(f) >> 16);
    }
  }
}

void slice_rows_bf16(int S, const double* src, int64_t ld, int rows, int K,
                     uint16_t* planes, int* expo, hipStream_t s) {
  if (rows <= 0) return;
  const int64_t Kp = slice_kpad(K);
  switch (S) {
    case 2:
      hipLaunchKernelGGL(k_slice_rows<2>, dim3(rows), dim3(256), 0, s, src,
                         ld, rows, K, Kp, planes, expo);
      break;
    case 3:
      hipLaunchKernelGGL(k_slice_rows<3>, dim3(rows), dim3(256), 0, s, src,
                         ld, rows, K, Kp, planes, expo);
      break;
    case 4:
      hipLaunchKernelGGL(k_slice_rows<4>, dim3(rows), dim3(256), 0, s, src,
                         ld, rows, K, Kp, planes, expo);
      break;
    default:
      fail(GAIN_ERR_ARG, "slice count %d not supported (2..4)", S);
  }
  check_launch("slice_rows");
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_transpose_mu(
    const double* __restrict__ mu, int nc, double* __restrict__ muT) {
  __shared__ double t[32][33];
  const int a = blockIdx.z;
  const int i0 = blockIdx.y * 32, f0 = blockIdx.x * 32;
  const double* in = mu + (int64_t)a * nc * nc;
  double* out = muT + (int64_t)a * nc * nc;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int y = ty; y < 32; y += 8) {
    const int vi = i0 + y, vf = f0 + tx;
    t[y][tx] = (vi < nc && vf < nc) ? in[(int64_t)vi * nc + vf] : 0.0;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int vf = f0 + y, vi = i0 + tx;
    if (vi < nc && vf < nc) out[(int64_t)vf * nc + vi] = t[tx][y];
  }
}

void transpose_mu(const double* mu, int nc, double* muT, hipStream_t s) {
  dim3 grid((nc + 31) / 32, (nc + 31) / 32, 3);
  hipLaunchKernelGGL(k_transpose_mu, grid, dim3(256), 0, s, mu, nc, muT);
  check_launch("transpose_mu");
}

// ---------------------------------------------------------------------------
// Workgroup tile 64 x 64, 4 waves as 2 x 2 (32 x 32 each = 2 x 2 MFMA
// tiles), K step 32.  Lane layout of v_mfma_f32_16x16x32_bf16: A(m, k) with
// m = lane%16, k = 8*(lane/16) + 0..7; B(k, n) with n = lane%16; output
// element e of lane l holds D(m = 4*(l/16) + e, n = l%16).
constexpr int kSBM = 64, kSBN = 64, kSBK = 32, kSRow = kSBK + 8;

template <int S>
__global__ __launch_bounds__(256) void k_gemm_sliced(SlicedGemmArgs g,
                                                     int ntn) {
  // stages per chunk (exact FP32 sums, one exponent pair per chunk)
  constexpr int kChunk = (S <= 3 ? 256 : 128) / kSBK;
  __shared__ __attribute__((aligned(16))) uint16_t As[S][kSBM][kSRow];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[S][kSBN][kSRow];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int nxy = ntn * gridDim.y;
  int lin = blockIdx.x + gridDim.x * blockIdx.y;
  if (nxy % 8 == 0) lin = (lin % 8) * (nxy / 8) + lin / 8;
  const int tn = lin % ntn, tm = lin / ntn;
  const int m0 = tm * kSBM, n0 = tn * kSBN;
  const int z = blockIdx.z;
  const uint16_t* A = g.A + (int64_t)z * g.a_z;
  const int* eA = g.eA + (int64_t)z * g.ea_z;
  double* C = g.C + (int64_t)z * g.c_z;

  f4 acc[S][2][2];
  double acc64[2][2][4];
#pragma unroll
  for (int L = 0; L < S; ++L)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[L][i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc64[i][j][e] = 0.0;

  const int lr = tid >> 2, lk = (tid & 3) * 8;
  const bool arow = (m0 + lr) < g.M, brow = (n0 + lr) < g.N;
  u4 pa[S], pb[S];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < S; ++i) {
      pa[i] = arow ? *(const u4*)(A + i * g.a_plane + (int64_t)(m0 + lr) * g.Kp +
                                  k0 + lk)
                   : u4{0u, 0u, 0u, 0u};
      pb[i] = brow ? *(const u4*)(g.B + i * g.b_plane +
                                  (int64_t)(n0 + lr) * g.Kp + k0 + lk)
                   : u4{0u, 0u, 0u, 0u};
    }
  };
  // rows / columns this lane's accumulators belong to (clamped in range)
  int mrow[2][4], ncol[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
      mrow[t][e] = min(m0 + wm * 32 + t * 16 + 4 * (lane >> 4) + e, g.M - 1);
    ncol[t] = min(n0 + wn * 32 + t * 16 + (lane & 15), g.N - 1);
  }
  auto flush = [&](int ch) {
    int ea[2][4], eb[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
      for (int e = 0; e < 4; ++e) ea[t][e] = eA[(int64_t)mrow[t][e] * g.nch + ch];
      eb[t] = g.eB[(int64_t)ncol[t] * g.nch + ch];
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          double sum = 0.0;
#pragma unroll
          for (int L = S - 1; L >= 0; --L) sum += (double)acc[L][i][j][e];
          acc64[i][j][e] += ldexp(sum, ea[i][e] + eb[j]);
        }
#pragma unroll
    for (int L = 0; L < S; ++L)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[L][i][j] = f4{0.f, 0.f, 0.f, 0.f};
  };

  load(0);
  int stage = 0, chunk = 0;
  for (int64_t k0 = 0; k0 < g.Kp; k0 += kSBK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < S; ++i) {
      *(u4*)&As[i][lr][lk] = pa[i];
      *(u4*)&Bs[i][lr][lk] = pb[i];
    }
    __syncthreads();
    if (k0 + kSBK < g.Kp) load(k0 + kSBK);
    bf8 fa[S][2], fb[S][2];
    const int fk = 8 * (lane >> 4);
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        fa[i][t] = __builtin_bit_cast(
            bf8, *(const u4*)&As[i][wm * 32 + t * 16 + (lane & 15)][fk]);
        fb[i][t] = __builtin_bit_cast(
            bf8, *(const u4*)&Bs[i][wn * 32 + t * 16 + (lane & 15)][fk]);
      }
#pragma unroll
    for (int L = 0; L < S; ++L)
#pragma unroll
      for (int i = 0; i <= L; ++i)
#pragma unroll
        for (int tm2 = 0; tm2 < 2; ++tm2)
#pragma unroll
          for (int tn2 = 0; tn2 < 2; ++tn2)
            acc[L][tm2][tn2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                fa[i][tm2], fb[L - i][tn2], acc[L][tm2][tn2], 0, 0, 0);
    if (++stage == kChunk) {
      stage = 0;
      flush(chunk++);
    }
  }
#pragma unroll
  for (int tm2 = 0; tm2 < 2; ++tm2)
#pragma unroll
    for (int tn2 = 0; tn2 < 2; ++tn2) {
      const int n = n0 + wn * 32 + tn2 * 16 + (lane & 15);
      if (n >= g.N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = m0 + wm * 32 + tm2 * 16 + 4 * (lane >> 4) + e;
        if (m < g.M) C[(int64_t)n * g.ldc + m] = acc64[tm2][tn2][e];
      }
    }
}

void gemm_sliced_bf16(const SlicedGemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  GAIN_CHECK_ARG(g.Kp % 256 == 0 && g.nch == g.Kp / slice_chunk(g.S),
                 "sliced gemm: Kp / chunk count mismatch");
  const int ntm = (g.M + kSBM - 1) / kSBM, ntn = (g.N + kSBN - 1) / kSBN;
  dim3 grid(ntn, ntm, g.nz);
  switch (g.S) {
    case 2:
      hipLaunchKernelGGL(k_gemm_sliced<2>, grid, dim3(256), 0, s, g, ntn);
      break;
    case 3:
      hipLaunchKernelGGL(k_gemm_sliced<3>, grid, dim3(256), 0, s, g, ntn);
      break;
    case 4:
      hipLaunchKernelGGL(k_gemm_sliced<4>, grid, dim3(256), 0, s, g, ntn);
      break;
    default:
      fail(GAIN_ERR_ARG, "slice count %d not supported (2..4)", g.S);
  }
  check_launch("gemm_sliced");
}

}  // namespace gain
