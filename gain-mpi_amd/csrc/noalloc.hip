// Allocation-free form of the plain FP64 drop-in (gain_plain_lines_noalloc).
//
// Same launches and arguments as gain_plain_lines for one initial state
// (scatter + expansion, mu X on the FP64 matrix cores, gather, transposed
// expansion, GEMV over the HBM-resident final records, epilogue), but every
// work buffer is carved out of one fixed device array, `scratch`, and the
// per-call index arrays reach the device as kernel arguments, so a call makes
// no device or pinned-host allocation, no memset and no host-to-device copy.
// Each GPU the library runs on holds its own copy of the array until the
// process ends; calls on one GPU share it, so calls that overlap on one GPU
// (other contexts, threads or streams) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine_internal.hpp"

namespace gain {
namespace {

constexpr size_t kScratchBytes = GAIN_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each GPU (hipGetSymbolAddress), filled at
// context creation with that GPU current.  The array's name is a host-side
// stand-in and never reaches a kernel.
constexpr int kMaxDevices = 64;
std::mutex g_lock;
unsigned char* g_scratch[kMaxDevices] = {};

// Per-call index arrays, passed by value and written into the scratch array
// by one small kernel.
constexpr int kWords = 64;
struct Words {
  int n;
  int w[kWords];
};

__global__ __launch_bounds__(64) void k_put_words(Words a, int* dst) {
  const int t = threadIdx.x;
  if (t < a.n) dst[t] = a.w[t];
}

// Bump allocator over the scratch array (256-byte aligned pieces).
struct Carve {
  unsigned char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t n) {
    off = (off + 255) / 256 * 256;
    T* p = (T*)(base + off);
    off += std::max<size_t>(n, 1) * sizeof(T);
    return p;
  }
};

}  // namespace

// Called by gain_create with `dev` current: resolves this GPU's copy of the
// scratch array once.
void noalloc_init_device(int dev) {
  GAIN_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d beyond %d", dev,
                 kMaxDevices);
  std::lock_guard<std::mutex> lk(g_lock);
  if (g_scratch[dev]) return;
  void* p = nullptr;
  GAIN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch[dev] = (unsigned char*)p;
}

// Plain FP64 lines of one initial state with every buffer in `scratch`.
static void plain_noalloc(gain_ctx* c, const Plan& P, const double* cI_dev,
                          const gain_lines* out) {
  Basis& bI = *P.bI;
  Basis& bF = *P.bF;
  Block& blkI = *P.blkI;
  Block& blkF = *P.blkF;
  const int nc = c->nc;
  if (bI.nvib > nc || bF.nvib > nc)
    fail(GAIN_ERR_ARG,
         "vibrational index of J=%d or J=%d exceeds the dipole dimension %d",
         bI.J, bF.J, nc);
  if (bI.slot_nc != nc)
    fail(GAIN_ERR_STATE, "scatter tables of J index %d are not built",
         P.jindI);
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_lock);
    base = g_scratch[c->device];
  }
  if (!base)
    fail(GAIN_ERR_STATE, "scratch array of device %d not resolved",
         c->device);
  const int nDI = (int)P.DI.size();
  const int nvpi = (int)P.vec.size();
  const int nvecX = nDI, nvecG = nvpi;  // one initial state
  const bool compute = P.Fhi > P.Flo && nvpi > 0 && nDI > 0;
  const double* cbase = cI_dev ? cI_dev : blkI.vec;
  if (compute && !cbase)
    fail(GAIN_ERR_STATE, "initial block has no eigenvectors and no cI given");

  // ---- layout: index words, X, Y, H (+ zero vector), G, GEMV partials
  Carve cv{base};
  int* words = cv.take<int>(kWords);
  const int64_t xstride = (int64_t)bI.ncol * nc;
  const int ncolTot = nvecX * bI.ncol;
  const int ks = mu_ksplit_auto(c, ncolTot);
  const int64_t hstride = bF.nprim;
  const Irrep& irF = bF.irrep[P.gF];
  const int64_t gstride = (irF.ntot + 1) / 2 * 2;
  const int nR = P.Fhi - P.Flo;
  const int gks = c->gemv_ksplit;
  double* X = cv.take<double>((size_t)nvecX * xstride);
  double* Y = cv.take<double>((size_t)ks * 3 * ncolTot * nc);
  double* H = cv.take<double>((size_t)(nvecX + 1) * hstride);
  double* G = cv.take<double>((size_t)nvecG * gstride);
  double* part = cv.take<double>((size_t)gks * nR * nvpi);
  const size_t cap = c->noalloc_cap > 0
                         ? std::min<size_t>(kScratchBytes, c->noalloc_cap)
                         : kScratchBytes;
  if (cv.off > cap)
    fail(GAIN_ERR_ARG,
         "allocation-free call needs %zu bytes of scratch, above its %zu "
         "(mu X split-K %d, GEMV split-K %d: lower \"mu_ksplit\" or "
         "\"gemv_ksplit\", or use gain_plain_lines)",
         cv.off, cap, ks, gks);

  // ---- index words (the Small arrays of one initial state)
  Words wd{};
  int nw = 0;
  auto put64 = [&](int64_t v) {
    std::memcpy(&wd.w[nw], &v, sizeof(v));
    nw += 2;
  };
  const int w_coff = 0;
  for (int k = 0; k < nvecX; ++k)
    put64(cI_dev ? 0 : (int64_t)P.levels[0] * blkI.ld);
  const int w_ideg = nw;
  for (int k = 0; k < nvecX; ++k) wd.w[nw++] = P.DI[k];
  const int w_hv = nw;
  for (int v = 0; v < nvpi; ++v) {
    int pos = -1;
    for (int k = 0; k < nDI; ++k)
      if (P.DI[k] == P.vec[v].first) pos = k;
    // a (dI,dF) pair whose dI has no half line strength reads the zero
    // vector after the last H vector
    wd.w[nw++] = pos >= 0 ? pos : nvecX;
  }
  const int w_dv = nw;
  for (int v = 0; v < nvpi; ++v) wd.w[nw++] = P.vec[v].second;
  const int w_sv = nw;
  for (int v = 0; v < nvpi; ++v)
    wd.w[nw++] = P.vec[v].first + P.vec[v].second * c->maxdeg;
  wd.w[nw++] = 0;
  if (nw & 1) ++nw;  // the energy is 8-byte aligned
  const int w_e = nw;
  put64(0);
  std::memcpy(&wd.w[w_e], &P.eI[0], sizeof(double));
  const int w_l = nw;
  wd.w[nw++] = P.levels[0];
  const int w_ok = nw;
  wd.w[nw++] = P.okI[0];  // little-endian: the low byte is the flag
  if (nw > kWords) fail(GAIN_ERR_ARG, "too many index words (%d)", nw);
  wd.n = nw;
  hipLaunchKernelGGL(k_put_words, dim3(1), dim3(64), 0, c->stream, wd, words);
  check_launch("put_words");
  Small sm{};
  sm.c_off = (int64_t*)(words + w_coff);
  sm.ideg = words + w_ideg;
  sm.h_of_vec = words + w_hv;
  sm.ideg_of_vec = words + w_dv;
  sm.slot_of_vec = words + w_sv;
  sm.eI = (double*)(words + w_e);
  sm.levelI = words + w_l;
  sm.okI = (uint8_t*)(words + w_ok);

  // ---- half line strength, transposed expansion, GEMV (FP64)
  const double* src = nullptr;
  if (compute) {
    launch_fill_zero(H + (size_t)nvecX * hstride, hstride, c->stream);
    launch_scatter_expand(make_scatter_args(c, P, cbase, sm, nvecX, X),
                          c->stream);
    if (c->use_mfma) {
      GemmArgs g = make_mu_gemm_args(c, X, ncolTot, Y, ks);
      g.cfg = gemm_cfg_lookup(c, 1, g);
      gemm_f64(g, c->stream);
    } else {
      launch_mu_gemm_valu(c->d_mu, nc, X, ncolTot, Y, ks, c->stream);
    }
    if (bI.rotsym && bF.slot_nc != nc)
      fail(GAIN_ERR_STATE, "scatter tables of J index %d are not built",
           P.jindF);
    launch_gather_for(c, P, nvecX, ks, Y, H, c->stream);
    launch_expand_T(make_expand_T_args(P, H, sm, nvecG, G, gstride),
                    c->stream);
    GemvArgs a{};
    a.C = blkF.vec;
    a.ld = blkF.ld;
    a.ntot = irF.ntot;
    a.F0 = P.Flo;
    a.nF = nR;
    a.G = G;
    a.gstride = gstride;
    a.nvec = nvpi;
    a.partial = part;
    a.ksplit = gks;
    a.rows = c->gemv_rows;
    launch_gemv(a, c->stream);
    src = part;
  }
  LinesArgs la = make_lines_args(c, P, sm, nvpi, compute);
  la.src = src;
  la.ksplit = gks;
  la.src_ks = (int64_t)nR * nvpi;
  la.src_f = nvpi;
  la.src_v = 1;
  la.ls = out->ls;
  la.S = out->S;
  la.A = out->A;
  la.I = out->I;
  la.flags = out->flags;
  launch_lines(la, c->stream);
}

}  // namespace gain

using namespace gain;

extern "C" int gain_plain_lines_noalloc(gain_ctx* c, int jindI, int gammaI,
                                        int levelI, const double* cI_device,
                                        int jindF, int gammaF,
                                        const gain_lines* out) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(out, "null output descriptor");
    GAIN_CHECK_ARG(out->on_device,
                   "allocation-free call writes device outputs only "
                   "(on_device = 1)");
    DeviceGuard g(c->device);
    Plan P = make_plan(c, jindI, gammaI, &levelI, 1, jindF, gammaF);
    plain_noalloc(c, P, cI_device, out);
  });
}
