// Plain FP64 path kernels for gfx950 (wave64).  See kernels.hpp for the HBM
// layout of every buffer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "kernels.hpp"
#include "ozaki.hpp"

namespace gain {

typedef double dbl2 __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------------------
// Expand + scatter (dipole_GPU.cu:859-889 fused with the K/tau grouping of
// the half line strength).  One thread per (column, v) slot of X and
// kScatterNV consecutive vectors:
//   X[vec][2K+tau][v] = sum_{r in slot} (-1)^{(K mod 3) tau + K} x[r],
//   x[r] = sum_{i<N[case]} c[t0+i] repres[(t0+i) + d*Ntot + lambda*sdeg*Ntot]
// The sign is the initial-side factor of dipole_GPU.cu:437-438.  The slot's
// primitives and their symmetrisation cases are loaded once for all the
// thread's vectors, so the dependent index loads are shared.
__device__ inline void scatter_body(const ScatterArgs& a, int bx, int b0) {
  const int s = bx * 256 + threadIdx.x;
  const bool live = s < a.nslot;
  const int nv = min(kScatterNV, a.nvec - b0);
  double acc[kScatterNV];
  const double* cv[kScatterNV];
  int64_t dof[kScatterNV];
#pragma unroll
  for (int i = 0; i < kScatterNV; ++i) {
    acc[i] = 0.0;
    const int b = b0 + min(i, nv - 1);
    cv[i] = a.cbase + a.c_off[b];
    dof[i] = (int64_t)a.ideg[b] * a.ntot;
  }
  const int q0 = live ? a.slot_ptr[s] : 0;
  const int q1 = live ? a.slot_ptr[s + 1] : 0;
  for (int q = q0; q < q1; ++q) {
    const int r = a.slot_prim[q];
    const int icase = a.icontr[r];
    const int lam = a.icontr[r + a.nprim];
    const int t0 = a.ijt[icase];
    const int ne = a.N[icase];
    const double* rp = a.repres + (int64_t)lam * a.sdeg * a.ntot + t0;
    const int K = a.kval[r], t = a.tau[r];
    const bool neg = (((K % 3) * t + K) & 1) != 0;
#pragma unroll
    for (int i = 0; i < kScatterNV; ++i) {
      const double* c = cv[i] + t0;
      const double* p = rp + dof[i];
      double x = 0.0;
      for (int k = 0; k < ne; ++k) x = fma(c[k], p[k], x);
      acc[i] += neg ? -x : x;
    }
  }
#pragma unroll
  for (int i = 0; i < kScatterNV; ++i)
    if (i < nv && live) a.X[(int64_t)(b0 + i) * a.xstride + s] = acc[i];
}

__global__ __launch_bounds__(256) void k_scatter_expand(ScatterArgs a) {
  scatter_body(a, blockIdx.x, blockIdx.y * kScatterNV);
}

void launch_scatter_expand(const ScatterArgs& a, hipStream_t s) {
  if (a.nslot == 0 || a.nvec == 0) return;
  dim3 grid((a.nslot + 255) / 256, (a.nvec + kScatterNV - 1) / kScatterNV);
  hipLaunchKernelGGL(k_scatter_expand, grid, dim3(256), 0, s, a);
  check_launch("scatter_expand");
}

// ---------------------------------------------------------------------------
// mu_alpha * X on the vector ALUs (FP64 FMA).  Workgroup: 64 rows (vF) x 64
// columns, 4 waves each owning 16 columns; X staged through LDS in 32-row
// slices.  Split-K over vI writes partials (summed by the gather kernel).
__global__ __launch_bounds__(256) void k_mu_gemm_valu(
    const double* __restrict__ mu, int nc, const double* __restrict__ X,
    int ncolTot, double* __restrict__ Ypart, int ksplit, int kc) {
  __shared__ double xs[32][65];
  const int tx = threadIdx.x & 63;
  const int cg = threadIdx.x >> 6;
  const int row = blockIdx.x * 64 + tx;
  const int j0 = blockIdx.y * 64;
  const int alpha = blockIdx.z % 3;
  const int ks = blockIdx.z / 3;
  const int kb = ks * kc;
  const int ke = min(nc, kb + kc);
  const double* mua = mu + (int64_t)alpha * nc * nc;
  const int rrow = row < nc ? row : nc - 1;
  double acc[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = 0.0;
  for (int k0 = kb; k0 < ke; k0 += 32) {
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * 64; e += 256) {
      const int k = e & 31, j = e >> 5;
      const int gk = k0 + k, gj = j0 + j;
      xs[k][j] = (gk < ke && gj < ncolTot) ? X[(int64_t)gj * nc + gk] : 0.0;
    }
    __syncthreads();
    const int kn = min(32, ke - k0);
    for (int k = 0; k < kn; ++k) {
      const double m = mua[(int64_t)(k0 + k) * nc + rrow];
#pragma unroll
      for (int c = 0; c < 16; ++c) acc[c] = fma(m, xs[k][cg * 16 + c], acc[c]);
    }
  }
  if (row >= nc) return;
  double* yo = Ypart + ((int64_t)ks * 3 + alpha) * ncolTot * (int64_t)nc;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const int j = j0 + cg * 16 + c;
    if (j < ncolTot) yo[(int64_t)j * nc + row] = acc[c];
  }
  (void)ksplit;
}

void launch_mu_gemm_valu(const double* mu, int nc, const double* X,
                         int ncolTot, double* Ypart, int ksplit,
                         hipStream_t s) {
  int kc = (nc + ksplit - 1) / ksplit;
  kc = (kc + 31) / 32 * 32;
  dim3 grid((nc + 63) / 64, (ncolTot + 63) / 64, 3 * ksplit);
  hipLaunchKernelGGL(k_mu_gemm_valu, grid, dim3(256), 0, s, mu, nc, X,
                     ncolTot, Ypart, ksplit, kc);
  check_launch("mu_gemm_valu");
}

// ---------------------------------------------------------------------------
// Gather of the half line strength from mu_alpha X (dipole_GPU.cu:430-511):
//   h[r'] = (-1)^{(K_F mod 3) tau_F} sum_{K_I = K_F-1..K_F+1, K_I<=J_I}
//           f3j * s * sum_{tau_I} m * Y_alpha[v(r'), 2 K_I + tau_I]
// with alpha, m and s as in the reference kernel (s = truncated SQRT2 when
// K_I != K_F and K_I K_F != 0).
// Each thread forms one primitive of NB consecutive vectors: the
// primitive's K, tau and vibrational index and the rotational factors are
// formed once, and the up to 6 NB loads of Y (K_I = K-1..K+1, tau_I = 0, 1,
// NB vectors) are all independent, so a thread keeps many loads in flight.
template <int NB>
__device__ inline void gather_body(const GatherArgs& a, int bx, int b0) {
  const int r = bx * 256 + threadIdx.x;
  if (r >= a.nprimF) return;
  const int K = a.kvalF[r];
  const int tF = a.tauF[r] & 1;
  const int v = a.vibF[r];
  const int n1 = a.jmax + 1;
  double coef[3][2];
  int alpha[3][2], kcol[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int KI = K - 1 + q;
    const bool in = KI >= 0 && KI <= a.jI;
    kcol[q] = in ? KI : 0;
    const int dk = K - KI;
    const double f3j =
        in ? a.threej[a.jI + KI * n1 + (a.jF - a.jI + 1) * n1 * n1 +
                      (dk + 1) * n1 * n1 * 3]
           : 0.0;
    const double sq = (dk != 0 && KI * K != 0) ? kSqrt2Ref : 1.0;
#pragma unroll
    for (int tI = 0; tI < 2; ++tI) {
      int al, m;
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
      alpha[q][tI] = al;
      coef[q][tI] = f3j * sq * (double)m;
    }
  }
  double acc[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) acc[i] = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int tI = 0; tI < 2; ++tI) {
      if (coef[q][tI] == 0.0) continue;
      const double* yp = a.Y + alpha[q][tI] * a.ys_alpha +
                         (2 * kcol[q] + tI) * (int64_t)a.nc + v;
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int b = min(b0 + i, a.nvec - 1);
        const double* p = yp + (int64_t)b * a.ncolI * a.nc;
        double y = 0.0;
        for (int ks = 0; ks < a.ksplit; ++ks) y += p[ks * a.ys_ks];
        acc[i] = fma(coef[q][tI], y, acc[i]);
      }
    }
  const bool neg = ((K % 3) * tF) & 1;
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (b0 + i < a.nvec)
      a.H[(int64_t)(b0 + i) * a.hstride + r] = neg ? -acc[i] : acc[i];
}

constexpr int kGatherNB = 4;  // vectors per thread

__global__ __launch_bounds__(256) void k_gather_h(GatherArgs a) {
  gather_body<kGatherNB>(a, blockIdx.x, blockIdx.y * kGatherNB);
}

void launch_gather_h(const GatherArgs& a, hipStream_t s) {
  if (a.nprimF == 0 || a.nvec == 0) return;
  dim3 grid((a.nprimF + 255) / 256, (a.nvec + kGatherNB - 1) / kGatherNB);
  hipLaunchKernelGGL(k_gather_h, grid, dim3(256), 0, s, a);
  check_launch("gather_h");
}

// ---------------------------------------------------------------------------
// Symmetric-top gather: one thread per (v_F, vector) and kRotCF final
// columns.  The Wigner factors are uniform across the wave (scalar loads);
// the Y loads of consecutive v_F are coalesced.
constexpr int kRotCF = 4;

__global__ __launch_bounds__(64) void k_gather_rotsym(GatherRotArgs a) {
  const int v = blockIdx.x * 64 + threadIdx.x;
  const int c0 = blockIdx.y * kRotCF;
  const int b = blockIdx.z;
  if (v >= a.nc) return;
  const int64_t sF = 3ll * a.nlI, sdI = sF * a.nlF, sdF = sdI * a.md;
  int lF[kRotCF], dF[kRotCF];
#pragma unroll
  for (int q = 0; q < kRotCF; ++q) {
    const int cF = min(c0 + q, a.ncolF - 1);
    lF[q] = cF / a.md;
    dF[q] = cF - lF[q] * a.md;
  }
  double acc[kRotCF];
#pragma unroll
  for (int q = 0; q < kRotCF; ++q) acc[q] = 0.0;
  const double* yb = a.Y + (int64_t)b * a.ncolI * a.nc + v;
  for (int cI = 0; cI < a.ncolI; ++cI) {
    const int lI = cI / a.md, dI = cI - lI * a.md;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      const double* p = yb + n * a.ys_alpha + (int64_t)cI * a.nc;
      double y = 0.0;
      for (int ks = 0; ks < a.ksplit; ++ks) y += p[ks * a.ys_ks];
      const double* w = a.W + n + 3 * lI + sdI * dI;
#pragma unroll
      for (int q = 0; q < kRotCF; ++q)
        acc[q] = fma(w[sF * lF[q] + sdF * dF[q]], y, acc[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < kRotCF; ++q) {
    const int cF = c0 + q;
    if (cF >= a.ncolF) break;
    const int slot = cF * a.nc + v;
    for (int t = a.slot_ptrF[slot]; t < a.slot_ptrF[slot + 1]; ++t)
      a.H[(int64_t)b * a.hstride + a.slot_primF[t]] = acc[q];
  }
}

void launch_gather_rotsym(const GatherRotArgs& a, hipStream_t s) {
  if (a.nvec == 0 || a.ncolF == 0 || a.nc == 0) return;
  GAIN_CHECK_ARG(a.nvec < 65536, "rotsym gather: %d vectors", a.nvec);
  dim3 grid((a.nc + 63) / 64, (a.ncolF + kRotCF - 1) / kRotCF, a.nvec);
  hipLaunchKernelGGL(k_gather_rotsym, grid, dim3(64), 0, s, a);
  check_launch("gather_rotsym");
}

// ---------------------------------------------------------------------------
// Transpose of the expansion: g[t] = sum_{r in case(t)}
//   repres[t + d*Ntot + lambda(r)*sdeg*Ntot] h[r]
// so that <P c_F, h> = <c_F, g> for every final record c_F
// (GpuManager.cpp:733-741 expands each c_F and dots it with h instead).
__device__ inline void expand_T_body(const ExpandTArgs& a, int bx, int v) {
  const int t = bx * 256 + threadIdx.x;
  if (t >= a.ntot) return;
  const int c = a.tcase[t];
  const int d = a.ideg_of_vec[v];
  const double* h = a.H + (int64_t)a.h_of_vec[v] * a.hstride;
  const double* rp = a.repres + t + (int64_t)d * a.ntot;
  double acc = 0.0;
  for (int q = a.case_ptr[c]; q < a.case_ptr[c + 1]; ++q) {
    const int r = a.case_prim[q];
    const int lam = a.icontr[r + a.nprim];
    acc = fma(rp[(int64_t)lam * a.sdeg * a.ntot], h[r], acc);
  }
  a.G[(int64_t)v * a.gstride + t] = acc;
}

// Same for NV consecutive output vectors per thread: the case's primitive
// chain (tcase -> case_ptr -> case_prim -> icontr) is loaded once for all of
// them and their H loads are independent; each vector's sum is formed in
// the order of expand_T_body, so G is identical.
template <int NV>
__device__ inline void expand_T_nv(const ExpandTArgs& a, int bx, int v0) {
  const int t = bx * 256 + threadIdx.x;
  if (t >= a.ntot) return;
  const int nv = min(NV, a.nvec - v0);
  const int c = a.tcase[t];
  const double* h[NV];
  const double* rp[NV];
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const int v = v0 + min(i, nv - 1);
    h[i] = a.H + (int64_t)a.h_of_vec[v] * a.hstride;
    rp[i] = a.repres + t + (int64_t)a.ideg_of_vec[v] * a.ntot;
    acc[i] = 0.0;
  }
  const int64_t lstride = (int64_t)a.sdeg * a.ntot;
  for (int q = a.case_ptr[c]; q < a.case_ptr[c + 1]; ++q) {
    const int r = a.case_prim[q];
    const int64_t off = (int64_t)a.icontr[r + a.nprim] * lstride;
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = fma(rp[i][off], h[i][r], acc[i]);
  }
#pragma unroll
  for (int i = 0; i < NV; ++i)
    if (i < nv) a.G[(int64_t)(v0 + i) * a.gstride + t] = acc[i];
}

constexpr int kExpandNV = 4;  // output vectors per thread

__global__ __launch_bounds__(256) void k_expand_T(ExpandTArgs a) {
  expand_T_nv<kExpandNV>(a, blockIdx.x, blockIdx.y * kExpandNV);
}

void launch_expand_T(const ExpandTArgs& a, hipStream_t s) {
  if (a.ntot == 0 || a.nvec == 0) return;
  dim3 grid((a.ntot + 255) / 256, (a.nvec + kExpandNV - 1) / kExpandNV);
  hipLaunchKernelGGL(k_expand_T, grid, dim3(256), 0, s, a);
  check_launch("expand_T");
}

// ---------------------------------------------------------------------------
// FP64 GEMV over HBM-resident eigenvector records (the plain drop-in
// projection).  One wave per (group of R rows, K chunk); each lane streams
// 16-byte non-temporal loads of R rows while the NV projection vectors come
// from L2.  Wave-level tree reduction; split-K partials reduced by the
// epilogue in a fixed order.
template <int R, int NV, bool VEC>
__device__ inline void gemv_body(const GemvArgs& a, int kchunk,
                                 int64_t wave) {
  const int lane = threadIdx.x & 63;
  const int ngroups = (a.nF + R - 1) / R;
  if (wave >= (int64_t)ngroups * a.ksplit) return;
  const int grp = (int)(wave / a.ksplit);
  const int ks = (int)(wave % a.ksplit);
  const int tbeg = ks * kchunk;
  const int tend = min(a.ntot, tbeg + kchunk);
  const double* rowp[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    int F = grp * R + i;
    if (F >= a.nF) F = a.nF - 1;
    rowp[i] = a.C + (int64_t)(a.F0 + F) * a.ld;
  }
  double acc[R][NV];
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[i][v] = 0.0;

  if (VEC) {
    int t = tbeg + 2 * lane;
    // two 1 KiB steps per iteration keep 2R loads in flight per lane
    for (; t + 129 < tend; t += 256) {
      dbl2 c0[R], c1[R];
#pragma unroll
      for (int i = 0; i < R; ++i) {
        c0[i] = __builtin_nontemporal_load((const dbl2*)(rowp[i] + t));
        c1[i] = __builtin_nontemporal_load((const dbl2*)(rowp[i] + t + 128));
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if (v >= a.nvec) continue;
        const dbl2 g0 = *(const dbl2*)(a.G + v * a.gstride + t);
        const dbl2 g1 = *(const dbl2*)(a.G + v * a.gstride + t + 128);
#pragma unroll
        for (int i = 0; i < R; ++i) {
          double s = acc[i][v];
          s = fma(c0[i].x, g0.x, s);
          s = fma(c0[i].y, g0.y, s);
          s = fma(c1[i].x, g1.x, s);
          s = fma(c1[i].y, g1.y, s);
          acc[i][v] = s;
        }
      }
    }
    for (; t + 1 < tend; t += 128) {
      dbl2 c0[R];
#pragma unroll
      for (int i = 0; i < R; ++i)
        c0[i] = __builtin_nontemporal_load((const dbl2*)(rowp[i] + t));
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if (v >= a.nvec) continue;
        const dbl2 g0 = *(const dbl2*)(a.G + v * a.gstride + t);
#pragma unroll
        for (int i = 0; i < R; ++i) {
          acc[i][v] = fma(c0[i].x, g0.x, acc[i][v]);
          acc[i][v] = fma(c0[i].y, g0.y, acc[i][v]);
        }
      }
    }
    if (t < tend) {  // odd-length tail: one element left for one lane
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if (v >= a.nvec) continue;
        const double g = a.G[v * a.gstride + t];
#pragma unroll
        for (int i = 0; i < R; ++i) acc[i][v] = fma(rowp[i][t], g, acc[i][v]);
      }
    }
  } else {
    for (int t = tbeg + lane; t < tend; t += 64) {
      double c[R];
#pragma unroll
      for (int i = 0; i < R; ++i) c[i] = __builtin_nontemporal_load(rowp[i] + t);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        if (v >= a.nvec) continue;
        const double g = a.G[v * a.gstride + t];
#pragma unroll
        for (int i = 0; i < R; ++i) acc[i][v] = fma(c[i], g, acc[i][v]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int F = grp * R + i;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      if (v >= a.nvec) continue;
      double x = acc[i][v];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
      if (lane == 0 && F < a.nF)
        a.partial[((int64_t)ks * a.nF + F) * a.nvec + v] = x;
    }
  }
}

template <int R, int NV, bool VEC>
__global__ __launch_bounds__(256) void k_gemv(GemvArgs a, int kchunk) {
  gemv_body<R, NV, VEC>(a, kchunk,
                        (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6));
}

int gemv_kchunk(int ntot, int ksplit) {
  int kc = (ntot + ksplit - 1) / ksplit;
  return (kc + 127) / 128 * 128;
}

int64_t gemv_grid_items(const GemvArgs& a) {
  const int R = a.rows == 2 || a.rows == 8 ? a.rows : 4;
  const int64_t waves = (int64_t)((a.nF + R - 1) / R) * a.ksplit;
  return (waves + 3) / 4 * 256;
}

template <int R, int NV>
static void gemv_dispatch_vec(const GemvArgs& a, bool vec, hipStream_t s) {
  const int kc = gemv_kchunk(a.ntot, a.ksplit);
  const int64_t waves = (int64_t)((a.nF + R - 1) / R) * a.ksplit;
  dim3 grid((unsigned)((waves + 3) / 4));
  if (vec)
    hipLaunchKernelGGL((k_gemv<R, NV, true>), grid, dim3(256), 0, s, a, kc);
  else
    hipLaunchKernelGGL((k_gemv<R, NV, false>), grid, dim3(256), 0, s, a, kc);
}

template <int R>
static void gemv_dispatch_nv(const GemvArgs& a, bool vec, hipStream_t s) {
  if (a.nvec <= 1)
    gemv_dispatch_vec<R, 1>(a, vec, s);
  else if (a.nvec <= 2)
    gemv_dispatch_vec<R, 2>(a, vec, s);
  else if (a.nvec <= 4)
    gemv_dispatch_vec<R, 4>(a, vec, s);
  else
    gemv_dispatch_vec<R, 9>(a, vec, s);
}

void launch_gemv(const GemvArgs& a, hipStream_t s) {
  if (a.nF <= 0) return;
  GAIN_CHECK_ARG(a.nvec >= 1 && a.nvec <= 9, "gemv: nvec %d out of range",
                 a.nvec);
  const bool vec = (a.ld % 2 == 0) && (a.gstride % 2 == 0) &&
                   ((uintptr_t)a.C % 16 == 0) && ((uintptr_t)a.G % 16 == 0);
  switch (a.rows) {
    case 2: gemv_dispatch_nv<2>(a, vec, s); break;
    case 8: gemv_dispatch_nv<8>(a, vec, s); break;
    default: gemv_dispatch_nv<4>(a, vec, s); break;
  }
  check_launch("gemv");
}

// ---------------------------------------------------------------------------
// Epilogue: transition filters (States.cpp:46-76), S from the degenerate
// components, Einstein A, threshold and intensity (Output.cpp:52-112).
__device__ inline void lines_body(const LinesArgs& a, int bx, int i) {
  const int F = bx * 256 + threadIdx.x;
  if (F >= a.nF) return;
  const int64_t o = (int64_t)i * a.nF + F;
  const double eI = a.eI[i];
  const double eF = a.eF[F];
  const double nu = eF - eI;
  const double eu = eF - a.zpe;
  const bool pass = a.block_ok && a.okI[i] && nu >= a.fw0 && nu <= a.fw1 &&
                    eu <= a.eu1 && eu >= a.eu0 && a.isym_do_F &&
                    !(a.same_block && F == a.levelI[i]);
  const bool inside = F >= a.Flo && F < a.Fhi;
  const int md2 = a.maxdeg * a.maxdeg;
  double lsv[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) lsv[q] = 0.0;
  double ssum = 0.0;
  if (pass && inside) {
    for (int v = 0; v < a.nvec_per_i; ++v) {
      const double* p = a.src + (int64_t)(F - a.Flo) * a.src_f +
                        ((int64_t)i * a.nvec_per_i + v) * a.src_v;
      double val = 0.0;
      for (int ks = 0; ks < a.ksplit; ++ks) val += p[ks * a.src_ks];
      lsv[a.slot_of_vec[v]] = val;
      ssum += val * val;
    }
  }
  double S = 0.0, A = 0.0, In = 0.0;
  uint8_t fl = 0;
  if (pass) {
    fl = 1;
    if (!(nu < 1e-40)) {
      double s = ssum / (double)a.ndegI;
      if (a.reduced && a.ndegF != 1 && a.ndegI != 1) s *= (double)a.ndegI;
      A = kACoef * (double)(2 * a.jI + 1) * s * fabs(nu) * fabs(nu) *
          fabs(nu);
      S = s;
      if (!(A < a.threshold)) {
        fl |= 2;
        In = kCMCoef * A * a.gnsF * (2.0 * (double)a.jF + 1.0) *
             exp(-a.beta * (eI - a.zpe)) * (1.0 - exp(-a.beta * nu)) /
             (nu * nu * a.partition);
      }
    }
  }
  if (a.ls)
    for (int q = 0; q < md2; ++q) a.ls[o * md2 + q] = lsv[q];
  if (a.S) a.S[o] = S;
  if (a.A) a.A[o] = A;
  if (a.I) a.I[o] = In;
  if (a.flags) a.flags[o] = fl;
}

__global__ __launch_bounds__(256) void k_lines(LinesArgs a) {
  lines_body(a, blockIdx.x, blockIdx.y);
}

void launch_lines(const LinesArgs& a, hipStream_t s) {
  if (a.nF == 0 || a.nI == 0) return;
  dim3 grid((a.nF + 255) / 256, a.nI);
  hipLaunchKernelGGL(k_lines, grid, dim3(256), 0, s, a);
  check_launch("lines");
}

// ---------------------------------------------------------------------------
// FP64 TN GEMM on the vector ALUs: 64x64 tile, K step 16, 4x4 outputs per
// thread, both operands staged k-major in LDS.
__global__ __launch_bounds__(256) void k_gemm_tn_valu(
    int M, int N, int K, const double* __restrict__ A, int64_t lda,
    const double* __restrict__ B, int64_t ldb, double* __restrict__ C,
    int64_t ldc) {
  __shared__ double as[16][65];
  __shared__ double bs[16][65];
  const int tid = threadIdx.x;
  const int tm = tid & 15, tn = tid >> 4;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int k0 = 0; k0 < K; k0 += 16) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      const int k = e & 15, r = e >> 4;
      const int gk = k0 + k;
      const int gm = m0 + r, gn = n0 + r;
      as[k][r] = (gk < K && gm < M) ? A[(int64_t)gm * lda + gk] : 0.0;
      bs[k][r] = (gk < K && gn < N) ? B[(int64_t)gn * ldb + gk] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = as[k][tm + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = bs[k][tn + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n = n0 + tn + 16 * j;
    if (n >= N) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + tm + 16 * i;
      if (m < M) C[(int64_t)n * ldc + m] = acc[i][j];
    }
  }
}

void launch_gemm_tn_valu(int M, int N, int K, const double* A, int64_t lda,
                         const double* B, int64_t ldb, double* C,
                         int64_t ldc, hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((M + 63) / 64, (N + 63) / 64);
  hipLaunchKernelGGL(k_gemm_tn_valu, grid, dim3(256), 0, s, M, N, K, A, lda,
                     B, ldb, C, ldc);
  check_launch("gemm_tn_valu");
}

}  // namespace gain

namespace gain {

// ---------------------------------------------------------------------------
// Adjoint kernels.
__global__ __launch_bounds__(256) void k_dls(DlsArgs a) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (f >= a.nR) return;
  const int i = v / a.nvpi;
  const int F = a.Flo + f;
  const int64_t o = (int64_t)i * a.nF + F;
  // the forward leaves S = 0 for nu < 1e-40 (Output.cpp:65)
  const double nu = a.eF[F] - a.eI[i];
  const double g = ((a.flags[o] & 1) && !(nu < 1e-40)) ? a.dS[o] : 0.0;
  const int64_t p = (int64_t)v * a.nR + f;
  const double ls =
      a.ls_pair ? a.ls_pair[o * a.md2 + a.slot_of_vec[v - i * a.nvpi]]
                : a.ls[p];
  a.dls[p] = g * a.scale * ls;
}

void launch_dls(const DlsArgs& a, hipStream_t s) {
  if (a.nR <= 0) return;
  dim3 grid((a.nR + 255) / 256, a.nI * a.nvpi);
  hipLaunchKernelGGL(k_dls, grid, dim3(256), 0, s, a);
  check_launch("dls");
}

__global__ __launch_bounds__(256) void k_expand_acc(ExpandAccArgs a) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int hv = blockIdx.y;
  if (r >= a.nprim) return;
  const int icase = a.icontr[r];
  const int lam = a.icontr[r + a.nprim];
  const int t0 = a.ijt[icase];
  const int ne = a.N[icase];
  double acc = 0.0;
  for (int q = a.vptr[hv]; q < a.vptr[hv + 1]; ++q) {
    const double* g = a.dG + (int64_t)a.vidx[q] * a.gstride;
    const double* rp = a.repres + (int64_t)a.vdeg[q] * a.ntot +
                       (int64_t)lam * a.sdeg * a.ntot;
    for (int k = 0; k < ne; ++k) acc = fma(g[t0 + k], rp[t0 + k], acc);
  }
  a.dH[(int64_t)hv * a.hstride + r] = acc;
}

void launch_expand_acc(const ExpandAccArgs& a, hipStream_t s) {
  if (a.nprim <= 0 || a.nH <= 0) return;
  dim3 grid((a.nprim + 255) / 256, a.nH);
  hipLaunchKernelGGL(k_expand_acc, grid, dim3(256), 0, s, a);
  check_launch("expand_acc");
}

// dY[alpha][b, 2KI+tI][v] = sum_{KF in KI-1..KI+1, tF} [link alpha]
//   f3j * s * m * (-1)^{(KF mod 3) tF} * sum_{r' in slot(KF, tF, v)} dH[r']
__global__ __launch_bounds__(256) void k_gather_T(GatherTArgs a) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const int colI = blockIdx.y;  // 2 KI + tI
  const int b = blockIdx.z;
  if (v >= a.nc) return;
  const int KI = colI >> 1, tI = colI & 1;
  const int n1 = a.jmax + 1;
  const double* dh = a.dH + (int64_t)b * a.hstride;
  double acc[3] = {0.0, 0.0, 0.0};
  if (KI <= a.jI) {
    for (int KF = max(KI - 1, 0); KF <= min(KI + 1, a.jF); ++KF) {
      const int dk = KF - KI;
      const double f3j = a.threej[a.jI + KI * n1 + (a.jF - a.jI + 1) * n1 * n1 +
                                  (dk + 1) * n1 * n1 * 3];
      const double sq = (dk != 0 && KI * KF != 0) ? kSqrt2Ref : 1.0;
      for (int tF = 0; tF < 2; ++tF) {
        int alpha, m;
        if (dk == 0) {
          alpha = 2;
          m = tF - tI;
        } else if (tF != tI) {
          alpha = 0;
          m = (tF - tI) * dk;
        } else {
          alpha = 1;
          m = -1;
        }
        if (m == 0) continue;
        const int slot = (2 * KF + tF) * a.nc + v;
        double sum = 0.0;
        for (int q = a.slot_ptrF[slot]; q < a.slot_ptrF[slot + 1]; ++q)
          sum += dh[a.slot_primF[q]];
        const double sgn = (((KF % 3) * tF) & 1) ? -1.0 : 1.0;
        acc[alpha] = fma(f3j * sq * (double)m * sgn, sum, acc[alpha]);
      }
    }
  }
  const int64_t col = (int64_t)b * a.ncolI + colI;
  const int64_t zs = (int64_t)a.nvec * a.ncolI * a.nc;
#pragma unroll
  for (int al = 0; al < 3; ++al) a.dY[al * zs + col * a.nc + v] = acc[al];
}

void launch_gather_T(const GatherTArgs& a, hipStream_t s) {
  if (a.nc <= 0 || a.nvec <= 0) return;
  dim3 grid((a.nc + 255) / 256, a.ncolI, a.nvec);
  hipLaunchKernelGGL(k_gather_T, grid, dim3(256), 0, s, a);
  check_launch("gather_T");
}

__global__ __launch_bounds__(256) void k_transpose(const double* __restrict__ in,
                                                   int rows, int cols,
                                                   double* __restrict__ out) {
  __shared__ double t[32][33];
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int y = ty; y < 32; y += 8) {
    const int r = r0 + y, c = c0 + tx;
    t[y][tx] = (r < rows && c < cols) ? in[(int64_t)r * cols + c] : 0.0;
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int c = c0 + y, r = r0 + tx;
    if (r < rows && c < cols) out[(int64_t)c * rows + r] = t[tx][y];
  }
}

void launch_transpose(const double* in, int rows, int cols, double* out,
                      hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  dim3 grid((cols + 31) / 32, (rows + 31) / 32);
  hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, s, in, rows, cols, out);
  check_launch("transpose");
}

__global__ __launch_bounds__(256) void k_axpy1(const double* __restrict__ x,
                                               double* __restrict__ y,
                                               int64_t n, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = accumulate ? y[i] + x[i] : x[i];
}

void launch_axpy1(const double* x, double* y, int64_t n, int accumulate,
                  hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpy1, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     s, x, y, n, accumulate);
  check_launch("axpy1");
}

// mu[i] = sum_p c[p] M[p][i]: one pass over the P basis matrices, two
// elements per thread with 16-byte loads.
__global__ __launch_bounds__(256) void k_combine(int P, int64_t n,
                                                 const double* __restrict__ c,
                                                 const double* __restrict__ M,
                                                 double* __restrict__ mu) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 2;
  if (i >= n) return;
  double a0 = 0.0, a1 = 0.0;
  if (i + 1 < n && (n % 2 == 0)) {
    for (int p = 0; p < P; ++p) {
      const dbl2 m = __builtin_nontemporal_load((const dbl2*)(M + p * n + i));
      a0 = fma(c[p], m.x, a0);
      a1 = fma(c[p], m.y, a1);
    }
    mu[i] = a0;
    mu[i + 1] = a1;
  } else {
    for (int p = 0; p < P; ++p) a0 = fma(c[p], M[p * n + i], a0);
    mu[i] = a0;
    if (i + 1 < n) {
      for (int p = 0; p < P; ++p) a1 = fma(c[p], M[p * n + i + 1], a1);
      mu[i + 1] = a1;
    }
  }
}

void launch_combine(int P, int64_t n, const double* c, const double* M,
                    double* mu, hipStream_t s) {
  if (n <= 0) return;
  const int64_t threads = (n + 1) / 2;
  hipLaunchKernelGGL(k_combine, dim3((unsigned)((threads + 255) / 256)),
                     dim3(256), 0, s, P, n, c, M, mu);
  check_launch("combine");
}


// ---------------------------------------------------------------------------
// Grouped launches (mixed-J sweeps): one kernel over a table of argument
// blocks in device memory; entry t owns workgroups [start[t], start[t+1])
// of the flattened grid and runs the same body as its single launch.
namespace {

__device__ inline int find_entry(const int* __restrict__ start, int n,
                                 int b) {
  int lo = 0, hi = n - 1;  // last t with start[t] <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

struct ScatterOp {
  using Args = ScatterArgs;
  static __host__ __device__ int nbx(const Args& a) {
    return (a.nslot + 255) / 256;
  }
  static int64_t blocks(const Args& a) {
    return a.nslot == 0
               ? 0
               : (int64_t)nbx(a) * ((a.nvec + kScatterNV - 1) / kScatterNV);
  }
  static __device__ void run(const Args& a, int local) {
    const int n = nbx(a);
    scatter_body(a, local % n, (local / n) * kScatterNV);
  }
};
struct GatherOp {
  using Args = GatherArgs;
  static __host__ __device__ int nbx(const Args& a) {
    return (a.nprimF + 255) / 256;
  }
  static int64_t blocks(const Args& a) {
    return a.nprimF == 0
               ? 0
               : (int64_t)nbx(a) * ((a.nvec + kGatherNB - 1) / kGatherNB);
  }
  static __device__ void run(const Args& a, int local) {
    const int n = nbx(a);
    gather_body<kGatherNB>(a, local % n, (local / n) * kGatherNB);
  }
};
struct ExpandTOp {
  using Args = ExpandTArgs;
  static __host__ __device__ int nbx(const Args& a) {
    return (a.ntot + 255) / 256;
  }
  static int64_t blocks(const Args& a) {
    return a.ntot == 0 ? 0 : (int64_t)nbx(a) * a.nvec;
  }
  static __device__ void run(const Args& a, int local) {
    const int n = nbx(a);
    expand_T_body(a, local % n, local / n);
  }
};
struct LinesOp {
  using Args = LinesArgs;
  static __host__ __device__ int nbx(const Args& a) {
    return (a.nF + 255) / 256;
  }
  static int64_t blocks(const Args& a) {
    return a.nF == 0 ? 0 : (int64_t)nbx(a) * a.nI;
  }
  static __device__ void run(const Args& a, int local) {
    const int n = nbx(a);
    lines_body(a, local % n, local / n);
  }
};

template <class Op>
__global__ __launch_bounds__(256) void k_grouped(
    const typename Op::Args* __restrict__ args, const int* __restrict__ start,
    int n) {
  const int t = find_entry(start, n, blockIdx.x);
  Op::run(args[t], blockIdx.x - start[t]);
}

// GEMV entries: R rows per wave and NV vector slots for every entry; the
// K chunk of entry t is gemv_kchunk(ntot, ksplit) of that entry.
// Workgroup ids are remapped so that each XCD walks a contiguous range of
// entries: the projected vectors of the packs one XCD has in flight then
// fit its L2 (they are re-read once per group of R rows).
template <int R, int NV>
__global__ __launch_bounds__(256) void k_gemv_grouped(
    const GemvArgs* __restrict__ args, const int* __restrict__ start, int n) {
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int t = find_entry(start, n, b);
  const GemvArgs& a = args[t];
  int kc = (a.ntot + a.ksplit - 1) / a.ksplit;
  kc = (kc + 127) / 128 * 128;
  gemv_body<R, NV, true>(a, kc,
                         (int64_t)(b - start[t]) * 4 + (threadIdx.x >> 6));
}

template <class Op>
void grouped_table(const std::vector<typename Op::Args>& v,
                   std::vector<int>* start, std::vector<int>* keep) {
  int64_t total = 0;
  start->clear();
  keep->clear();
  for (size_t t = 0; t < v.size(); ++t) {
    const int64_t b = Op::blocks(v[t]);
    if (b == 0) continue;
    keep->push_back((int)t);
    start->push_back((int)total);
    total += b;
  }
  GAIN_CHECK_ARG(total < (1ll << 31), "grouped launch: too many blocks");
  start->push_back((int)total);
}

}  // namespace

#define GAIN_GROUPED_TABLE(NAME, OP, ARGS)                                    \
  GroupedTable NAME(const std::vector<ARGS>& v) {                            \
    std::vector<int> start, keep;                                            \
    grouped_table<OP>(v, &start, &keep);                                     \
    GroupedTable g;                                                          \
    g.start = start;                                                         \
    g.keep = keep;                                                           \
    return g;                                                                \
  }
GAIN_GROUPED_TABLE(grouped_scatter_table, ScatterOp, ScatterArgs)
GAIN_GROUPED_TABLE(grouped_gather_table, GatherOp, GatherArgs)
GAIN_GROUPED_TABLE(grouped_expand_T_table, ExpandTOp, ExpandTArgs)
GAIN_GROUPED_TABLE(grouped_lines_table, LinesOp, LinesArgs)
#undef GAIN_GROUPED_TABLE

GroupedTable grouped_gemv_table(const std::vector<GemvArgs>& v, int R) {
  GroupedTable g;
  int64_t total = 0;
  for (size_t t = 0; t < v.size(); ++t) {
    const GemvArgs& a = v[t];
    if (a.nF <= 0) continue;
    GAIN_CHECK_ARG((a.ld % 2 == 0) && (a.gstride % 2 == 0) &&
                       ((uintptr_t)a.C % 16 == 0) &&
                       ((uintptr_t)a.G % 16 == 0),
                   "grouped gemv: records and vectors must be 16-byte "
                   "aligned with even strides");
    const int64_t waves = (int64_t)((a.nF + R - 1) / R) * a.ksplit;
    g.keep.push_back((int)t);
    g.start.push_back((int)total);
    total += (waves + 3) / 4;
  }
  GAIN_CHECK_ARG(total < (1ll << 31), "grouped gemv: too many blocks");
  g.start.push_back((int)total);
  return g;
}

void launch_grouped(int kind, const void* d_args, const int* d_start, int n,
                    int nblocks, hipStream_t s) {
  if (n <= 0 || nblocks <= 0) return;
  const dim3 grid((unsigned)nblocks), block(256);
  switch (kind) {
    case kGroupedScatter:
      hipLaunchKernelGGL(k_grouped<ScatterOp>, grid, block, 0, s,
                         (const ScatterArgs*)d_args, d_start, n);
      break;
    case kGroupedGather:
      hipLaunchKernelGGL(k_grouped<GatherOp>, grid, block, 0, s,
                         (const GatherArgs*)d_args, d_start, n);
      break;
    case kGroupedExpandT:
      hipLaunchKernelGGL(k_grouped<ExpandTOp>, grid, block, 0, s,
                         (const ExpandTArgs*)d_args, d_start, n);
      break;
    case kGroupedLines:
      hipLaunchKernelGGL(k_grouped<LinesOp>, grid, block, 0, s,
                         (const LinesArgs*)d_args, d_start, n);
      break;
    case kGroupedGemv4:
      hipLaunchKernelGGL((k_gemv_grouped<4, 4>), grid, block, 0, s,
                         (const GemvArgs*)d_args, d_start, n);
      break;
    case kGroupedGemv8:
      hipLaunchKernelGGL((k_gemv_grouped<8, 9>), grid, block, 0, s,
                         (const GemvArgs*)d_args, d_start, n);
      break;
    default:
      fail(GAIN_ERR_ARG, "grouped launch kind %d", kind);
  }
  check_launch("grouped");
}

}  // namespace gain
