// Kernel-level drop-ins for the free functions of GPU/dipole_GPU.h and the
// cuBLAS ddot of GpuManager.cpp:741.  Same argument meaning as the reference
// (device pointers, hipStream_t passed as void*).  The values the reference
// keeps in __constant__ memory (c_jmax, dip_stride_1, g_symnrepres,
// g_maxdegen; dipole_GPU.cu:67-71) are held host-side and passed to the
// kernels as arguments.
#include <hip/hip_runtime.h>

#include <atomic>

#include "common.hpp"
#include "kernels.hpp"

namespace gain {
namespace {
std::atomic<int> g_jmax{-1};
std::atomic<int> g_dip_stride{-1};
std::atomic<int> g_symnrepres{0};
std::atomic<int> g_maxdegen{0};

// device_expand_vectors semantics (dipole_GPU.cu:859-889)
__global__ __launch_bounds__(256) void k_expand(
    int dimen, int igamma, int maxcoeff, int ideg, int sdeg, int ntot,
    const int* __restrict__ ijterms, const int* __restrict__ icontr,
    const int* __restrict__ N, const double* __restrict__ repres,
    const double* __restrict__ vin, double* __restrict__ vout) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= dimen) return;
  const int irow = icontr[r];
  const int ib = icontr[r + dimen];
  const int t0 = ijterms[irow + igamma * maxcoeff];
  const int ne = N[irow];
  const double* rp =
      repres + (int64_t)ideg * ntot + (int64_t)ib * sdeg * ntot;
  double acc = 0.0;
  for (int i = 0; i < ne; ++i) acc = fma(vin[t0 + i], rp[t0 + i], acc);
  vout[r] = acc;
}

// One K_F block of the half line strength with the reference's argument
// meaning (dipole_GPU.cu:392-515): one thread per final primitive of the
// block; the initial K blocks K_F-1..K_F+1 are walked in 256-primitive
// tiles whose scaled coefficient, tau and vibrational row offset are staged
// in LDS.  Only entries whose vibrational index lies in the dipole piece
// [startFblock, endFblock) are written.
__global__ __launch_bounds__(256) void k_half_ls_kblock(
    int dimenF, int dimenI, int jI, int jF, int kF,
    const int* __restrict__ tauI, const int* __restrict__ tauF,
    const int* __restrict__ vibI, const int* __restrict__ vibF,
    int startF_idx, int startI_idx, int startFblock, int endFblock,
    int ncontrF, const int* __restrict__ kblockI,
    const double* __restrict__ dip, const double* __restrict__ x,
    const double* __restrict__ threej, int jmax, int dip_stride,
    double* __restrict__ half_ls) {
  __shared__ double s_coef[256];
  __shared__ int s_tau[256];
  __shared__ int s_row[256];
  const int t = threadIdx.x;
  const int rF = startF_idx + blockIdx.x * blockDim.x + t;
  const bool inblk = rF < startF_idx + dimenF;
  int vF = inblk ? vibF[rF] : startFblock - 1;
  const int tF = inblk ? (tauF[rF] & 1) : 0;
  const bool valid = inblk && vF >= startFblock && vF < endFblock;
  vF -= startFblock;
  const int64_t stride = (int64_t)ncontrF * dip_stride;
  const int n1 = jmax + 1;
  double acc = 0.0;
  int sI = startI_idx;
  for (int kI = max(kF - 1, 0); kI < kF + 2 && kI < jI + 1; ++kI) {
    const int kbs = kblockI[kI];
    const bool eq = (kI == kF);
    const double sq = (!eq && kI * kF != 0) ? kSqrt2Ref : 1.0;
    const double f3j = threej[jI + kI * n1 + (jF - jI + 1) * n1 * n1 +
                              (kF - kI + 1) * n1 * n1 * 3];
    for (int b0 = 0; b0 < kbs; b0 += 256) {
      __syncthreads();
      if (b0 + t < kbs) {
        const int rI = sI + b0 + t;
        const int ti = tauI[rI];
        const double sgn = ((((kI % 3) * ti) + kI) & 1) ? -1.0 : 1.0;
        s_coef[t] = x[rI] * sq * f3j * sgn;
        s_tau[t] = ti;
        s_row[t] = vibI[rI] * ncontrF;
      } else {
        s_coef[t] = 0.0;
        s_tau[t] = 0;
        s_row[t] = 0;
      }
      __syncthreads();
      if (valid) {
        const int n = min(256, kbs - b0);
        for (int j = 0; j < n; ++j) {
          const double c = s_coef[j];
          if (c == 0.0) continue;
          const int ti = s_tau[j];
          const bool tneq = (tF != ti);
          const int comp = eq ? 2 : (tneq ? 0 : 1);
          const int m = eq ? (tF - ti) : (tneq ? (tF - ti) * (kF - kI) : -1);
          acc = fma(c * (double)m, dip[vF + s_row[j] + comp * stride], acc);
        }
      }
    }
    sI += kbs;
    if (sI >= dimenI) break;
  }
  if (valid) {
    const bool neg = ((kF % 3) * tF) & 1;
    half_ls[rF] = neg ? -acc : acc;
  }
}

// Half line strength in the symmetric-top (ROTSYM) basis with the
// reference's argument meaning (dipole_GPU.cu:641-723).  A workgroup takes
// 64 final primitives; its four waves split every 256-wide tile of initial
// primitives (coefficient, dipole row and Wigner offset staged in LDS) into
// four 64-entry slices, and the four partial sums of each final primitive
// are added in LDS in a fixed order, so each entry is still assigned once
// (and a final block fills four times as many workgroups).  The
// 3-component Wigner factor is indexed n + 3 lvI + 3 nlvI lvF + 3 nlvI nlvF
// dgI + 3 nlvI nlvF maxdeg dgF and the dipole components lie
// ncontrF * dip_stride apart.  Only entries whose vibrational index lies in
// [startFblock, endFblock) are written.
__global__ __launch_bounds__(256) void k_half_ls_rotsym(
    int dimenF, int dimenI, int nlevelI, int nlevelF, int maxdeg,
    const int* __restrict__ degI, const int* __restrict__ degF,
    const int* __restrict__ levelI, const int* __restrict__ levelF,
    const int* __restrict__ vibI, const int* __restrict__ vibF,
    int startFblock, int endFblock, int ncontrF,
    const double* __restrict__ dip, const double* __restrict__ wigner,
    const double* __restrict__ x, int dip_stride,
    double* __restrict__ half_ls) {
  __shared__ double s_coef[256];
  __shared__ int64_t s_row[256];
  __shared__ int64_t s_w[256];
  __shared__ double s_part[4][64];
  const int t = threadIdx.x;
  const int f = t & 63, sl = t >> 6;  // final primitive, slice (= wave)
  const int rF = blockIdx.x * 64 + f;
  const bool inblk = rF < dimenF;
  int vF = inblk ? vibF[rF] : startFblock - 1;
  const bool valid = inblk && vF >= startFblock && vF < endFblock;
  vF -= startFblock;
  const int64_t comp = (int64_t)ncontrF * dip_stride;
  const int64_t sF = 3 * (int64_t)nlevelI, sdI = sF * nlevelF,
                sdF = sdI * maxdeg;
  const int64_t wF = valid ? sF * levelF[rF] + sdF * degF[rF] : 0;
  const double* dF = dip + (valid ? vF : 0);
  double acc = 0.0;
  for (int b0 = 0; b0 < dimenI; b0 += 256) {
    __syncthreads();
    const int rI = b0 + t;
    if (rI < dimenI) {
      s_coef[t] = x[rI];
      s_row[t] = (int64_t)vibI[rI] * ncontrF;
      s_w[t] = 3 * (int64_t)levelI[rI] + sdI * degI[rI];
    }
    __syncthreads();
    if (valid) {
      const int j1 = min(64 * sl + 64, dimenI - b0);
      for (int j = 64 * sl; j < j1; ++j) {
        const double* w = wigner + wF + s_w[j];
        const double* d = dF + s_row[j];
        double v = w[0] * d[0];
        v = fma(w[1], d[comp], v);
        v = fma(w[2], d[2 * comp], v);
        acc = fma(s_coef[j], v, acc);
      }
    }
  }
  s_part[sl][f] = acc;
  __syncthreads();
  if (sl == 0 && valid)
    half_ls[rF] = ((s_part[0][f] + s_part[1][f]) + s_part[2][f]) +
                  s_part[3][f];
}

// Deterministic single-workgroup dot product (fixed reduction order).
__global__ __launch_bounds__(1024) void k_ddot(int n,
                                               const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               double* __restrict__ out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) acc = fma(x[i], y[i], acc);
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x < 64) {
    double v = threadIdx.x < 16 ? red[threadIdx.x] : 0.0;
    for (int off = 8; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if (threadIdx.x == 0) *out = v;
  }
}
}  // namespace
}  // namespace gain

using namespace gain;

extern "C" int gain_copy_symmetry_constants(int sym_nrepres, int maxdeg) {
  return guarded([&] {
    GAIN_CHECK_ARG(sym_nrepres > 0 && sym_nrepres <= GAIN_MAX_IRREPS &&
                       maxdeg > 0 && maxdeg <= 3,
                   "copy_symmetry_constants: bad sizes %d %d", sym_nrepres,
                   maxdeg);
    g_symnrepres = sym_nrepres;
    g_maxdegen = maxdeg;
  });
}

extern "C" int gain_copy_jmax_constant(int jmax) {
  return guarded([&] {
    GAIN_CHECK_ARG(jmax >= 0, "copy_jmax_constant: jmax < 0");
    g_jmax = jmax;
  });
}

extern "C" int gain_copy_dipole_constant(int dip_stride) {
  return guarded([&] {
    GAIN_CHECK_ARG(dip_stride > 0, "copy_dipole_constant: stride <= 0");
    g_dip_stride = dip_stride;
  });
}

extern "C" int gain_transform_vector_primitive(
    int dimenI, int igammaI, int maxcoeff, int idegI, int sdeg, int Ntot,
    const int* ijterms, const int* icontr, const int* N,
    const double* repres, const double* vecI, double* vec, void* stream) {
  return guarded([&] {
    GAIN_CHECK_ARG(dimenI >= 0 && igammaI >= 0 && maxcoeff > 0 &&
                       idegI >= 0 && idegI < sdeg && Ntot >= 0,
                   "transform_vector_primitive: bad sizes");
    GAIN_CHECK_ARG(ijterms && icontr && N && repres && vecI && vec,
                   "transform_vector_primitive: null pointer");
    if (dimenI == 0) return;
    hipLaunchKernelGGL(k_expand, dim3((dimenI + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, dimenI, igammaI, maxcoeff, idegI,
                       sdeg, Ntot, ijterms, icontr, N, repres, vecI, vec);
    check_launch("expand");
  });
}

extern "C" int gain_compute_half_linestrength(
    int dimenF, int dimenI, int jI, int jF, int kF, const int* tauI,
    const int* tauF, const int* icorrI, const int* icorrF, int startF_idx,
    int startI_idx, int startFblock, int endFblock, int ncontrF,
    int kFBlocksize, const int* kblock_size, const double* dipole_me,
    const double* vector, const double* threej, double* half_ls,
    void* stream) {
  return guarded([&] {
    (void)kFBlocksize;
    const int jmax = g_jmax, stride = g_dip_stride;
    if (jmax < 0 || stride <= 0)
      fail(GAIN_ERR_STATE,
           "compute_half_linestrength: call gain_copy_jmax_constant and "
           "gain_copy_dipole_constant first");
    GAIN_CHECK_ARG(jI >= 0 && jF >= 0 && jI <= jmax && jF <= jmax &&
                       abs(jF - jI) <= 1 && kF >= 0 && kF <= jF,
                   "compute_half_linestrength: bad J/K (%d,%d,%d)", jI, jF,
                   kF);
    GAIN_CHECK_ARG(ncontrF == endFblock - startFblock && ncontrF > 0,
                   "compute_half_linestrength: ncontrF != endF-startF");
    GAIN_CHECK_ARG(tauI && tauF && icorrI && icorrF && kblock_size &&
                       dipole_me && vector && threej && half_ls,
                   "compute_half_linestrength: null pointer");
    if (dimenF <= 0) return;
    hipLaunchKernelGGL(k_half_ls_kblock, dim3((dimenF + 255) / 256),
                       dim3(256), 0, (hipStream_t)stream, dimenF, dimenI, jI,
                       jF, kF, tauI, tauF, icorrI, icorrF, startF_idx,
                       startI_idx, startFblock, endFblock, ncontrF,
                       kblock_size, dipole_me, vector, threej, jmax, stride,
                       half_ls);
    check_launch("half_ls_kblock");
  });
}

extern "C" int gain_compute_half_linestrength_rotsym(
    int dimenF, int dimenI, int nlevelI, int nlevelF, int sym_max_degen,
    const int* kI, const int* kF, const int* ktauI, const int* ktauF,
    const int* icorrI, const int* icorrF, int startFblock, int endFblock,
    int ncontrF, const double* dipole_me, const double* wigner,
    const double* vector, double* half_ls, void* stream) {
  return guarded([&] {
    const int stride = g_dip_stride;
    if (stride <= 0)
      fail(GAIN_ERR_STATE,
           "compute_half_linestrength_rotsym: call "
           "gain_copy_dipole_constant first");
    GAIN_CHECK_ARG(dimenF >= 0 && dimenI >= 0 && nlevelI > 0 &&
                       nlevelF > 0 && sym_max_degen > 0 &&
                       sym_max_degen <= 3,
                   "compute_half_linestrength_rotsym: bad sizes");
    GAIN_CHECK_ARG(ncontrF == endFblock - startFblock && ncontrF > 0,
                   "compute_half_linestrength_rotsym: ncontrF != endF-startF");
    GAIN_CHECK_ARG(kI && kF && ktauI && ktauF && icorrI && icorrF &&
                       dipole_me && wigner && vector && half_ls,
                   "compute_half_linestrength_rotsym: null pointer");
    if (dimenF == 0) return;
    hipLaunchKernelGGL(k_half_ls_rotsym, dim3((dimenF + 63) / 64),
                       dim3(256), 0, (hipStream_t)stream, dimenF, dimenI,
                       nlevelI, nlevelF, sym_max_degen, kI, kF, ktauI, ktauF,
                       icorrI, icorrF, startFblock, endFblock, ncontrF,
                       dipole_me, wigner, vector, stride, half_ls);
    check_launch("half_ls_rotsym");
  });
}

extern "C" int gain_ddot(int n, const double* x, const double* y,
                         double* result, void* stream) {
  return guarded([&] {
    GAIN_CHECK_ARG(n >= 0 && x && y && result, "ddot: bad arguments");
    hipLaunchKernelGGL(k_ddot, dim3(1), dim3(1024), 0, (hipStream_t)stream,
                       n, x, y, result);
    check_launch("ddot");
  });
}
