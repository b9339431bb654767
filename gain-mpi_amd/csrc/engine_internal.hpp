// Engine internals shared by the C-ABI translation units (engine.hip,
// sweep.hip): device buffers of one context, plans of one call, and the
// helpers that build them.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "kernels.hpp"
#include "ozaki.hpp"

namespace gain {

// Wigner 3j symbol (j1 j2 j3; m1 m2 m3), Racah formula in long double.
double three_j(int j1, int j2, int j3, int m1, int m2, int m3);


// ---------------------------------------------------------------------------
struct DeviceArena {
  std::vector<void*> ptrs;
  template <class T>
  T* alloc(size_t n) {
    if (n == 0) n = 1;
    void* p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess)
      fail(GAIN_ERR_NOMEM, "hipMalloc of %zu bytes failed", n * sizeof(T));
    ptrs.push_back(p);
    return (T*)p;
  }
  template <class T>
  T* upload(const T* h, size_t n) {
    T* d = alloc<T>(n);
    if (n) GAIN_HIP(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
  }
  void release(void* p) {
    auto it = std::find(ptrs.begin(), ptrs.end(), p);
    if (it != ptrs.end()) {
      (void)hipFree(*it);
      ptrs.erase(it);
    }
  }
  ~DeviceArena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
};

// Grow-only device work buffer.
struct Work {
  void* p = nullptr;
  size_t bytes = 0;
  template <class T>
  T* get(size_t n) {
    const size_t need = std::max<size_t>(n, 1) * sizeof(T);
    if (need > bytes) {
      if (p) {
        // earlier launches may still read the old buffer
        (void)hipDeviceSynchronize();
        (void)hipFree(p);
      }
      p = nullptr;
      size_t nb = std::max(need, bytes * 3 / 2);
      if (hipMalloc(&p, nb) != hipSuccess) {
        bytes = 0;
        fail(GAIN_ERR_NOMEM, "work buffer allocation of %zu bytes failed", nb);
      }
      bytes = nb;
    }
    return (T*)p;
  }
  ~Work() {
    if (p) (void)hipFree(p);
  }
};

// Pinned host staging ring for per-call parameter arrays: a slot is reused
// only after the copy that read it has executed.
struct PinnedRing {
  static constexpr int kSlots = 4;
  void* h[kSlots] = {};
  size_t bytes[kSlots] = {};
  hipEvent_t ev[kSlots] = {};
  bool used[kSlots] = {};
  int next = 0;
  void* acquire(size_t n, int* slot) {
    const int s = next;
    next = (next + 1) % kSlots;
    if (used[s]) GAIN_HIP(hipEventSynchronize(ev[s]));
    if (n > bytes[s]) {
      if (h[s]) (void)hipHostFree(h[s]);
      h[s] = nullptr;
      size_t nb = std::max<size_t>(n, 1 << 16);
      GAIN_HIP(hipHostMalloc(&h[s], nb, hipHostMallocDefault));
      bytes[s] = nb;
    }
    if (!ev[s]) GAIN_HIP(hipEventCreateWithFlags(&ev[s], hipEventDisableTiming));
    *slot = s;
    return h[s];
  }
  void release(int s, hipStream_t st) {
    GAIN_HIP(hipEventRecord(ev[s], st));
    used[s] = true;
  }
  ~PinnedRing() {
    for (int s = 0; s < kSlots; ++s) {
      if (ev[s]) {
        (void)hipEventSynchronize(ev[s]);
        (void)hipEventDestroy(ev[s]);
      }
      if (h[s]) (void)hipHostFree(h[s]);
    }
  }
};

struct Irrep {
  int ntot = 0, sdeg = 0;
  int* N = nullptr;
  double* repres = nullptr;
  int* tcase = nullptr;
  int maxN = 0;  // largest term count of a symmetrisation case
};

struct Basis {
  bool present = false, has_infl = false;
  int J = 0, nprim = 0, ncol = 0, nvib = 0;
  // symmetric-top (ROTSYM) basis: primitives carry a rotational level and a
  // degenerate component instead of (K, tau); X column = level * rmd + deg
  bool rotsym = false;
  int nlevels = 0, rmd = 0;
  std::vector<int> kstart, kblock, hk, htau, hvib;
  std::vector<int> hcol;  // X column of each primitive
  int *kval = nullptr, *tau = nullptr, *vib = nullptr;
  // scatter CSR over (col, v) slots, built for a given nc
  int slot_nc = -1;
  int *slot_ptr = nullptr, *slot_prim = nullptr;
  // per slot: its primitive when it holds exactly one, -1 when none, -2
  // when several (then the CSR lists them)
  int* slot_one = nullptr;
  bool slot_multi = false;  // some slot holds several primitives
  // inflation
  int maxsymcoeffs = 0, matsize = 0;
  int *ijterms = nullptr, *icontr = nullptr;
  int *case_ptr = nullptr, *case_prim = nullptr;
  Irrep irrep[GAIN_MAX_IRREPS];
};

// Sliced copy of a static operand (final records or the dipole): split-BF16
// planes with chunk exponents, or tiled split-INT8 planes (`i8`) with one
// exponent per row.
struct Sliced {
  int S = 0;           // slices (split forms) or moduli (CRT form)
  bool i8 = false;
  bool crt = false;    // int8 residues (CRT form), `beta` integer bits
  int beta = 0;
  int64_t Kp = 0;
  int64_t Rp = 0;      // padded rows of one (z) matrix (int8)
  int64_t plane = 0;   // bytes between int8 slice planes
  uint16_t* planes = nullptr;
  int8_t* p8 = nullptr;
  int* expo = nullptr;
  // CRT form of the dipole: nonzero k-block range of every row block of
  // each component (device and host copies, Rp / 32 entries per component)
  int2* kr = nullptr;
  std::vector<int2> kr_host;
};

struct Block {
  Sliced sl;
  int n = 0, rec_len = 0;
  int64_t ld = 0;
  const double* vec = nullptr;  // device, may be null (energies only)
  bool owned = false;
  std::vector<double> energy;
  double* d_energy = nullptr;
  bool sorted = true;
  std::vector<int> label;  // level number printed per state (default: row)
};

// Rotational factors of one (J_I, J_F) pair of symmetric-top bases
// (GpuManager::TransferWigner, GpuManager.cpp:440-470).
struct WignerTable {
  int nlI = 0, nlF = 0, md = 0;
  double* d = nullptr;  // [3][nlI][nlF][md][md], first index fastest
};

struct TimerSlot {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, pool;
  double total_ms = 0.0;
  int64_t launches = 0;
};

}  // namespace gain

// The opaque handle of the C ABI (global namespace, declared in
// gain_engine.h).
using namespace gain;

struct gain_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int jmax = 0, nrepres = 0, maxdeg = 0;
  std::vector<double> threej;
  double* d_threej = nullptr;
  gain_params params{};
  int pair[GAIN_MAX_IRREPS] = {};
  bool params_set = false;
  double* d_mu = nullptr;
  int nc = 0;
  // final vibrational columns of mu uploaded so far (pieces of
  // gain_upload_dipole); products need all nc of them
  std::vector<char> mu_have;
  int mu_missing = 0;
  Sliced mu_sl;
  std::vector<Basis> bases;
  std::map<std::pair<int, int>, Block> blocks;
  std::map<std::pair<int, int>, WignerTable> wigner;  // (jindI, jindF)
  DeviceArena arena;
  Work w_c, w_X, w_Y, w_H, w_G, w_part, w_small, w_out, w_tmp;
  Work w_xs, w_xe, w_gs, w_ge, w_res, w_desc;
  PinnedRing ring;
  // tuning
  int gemv_rows = 4, gemv_ksplit = 8, mu_ksplit = 0;  // 0: automatic
  int sweep_phases = 2;  // pipeline phases of a mixed-J sweep
  bool sweep_eager = false;  // replay sweeps by enqueueing (no graph)
  bool use_mfma = true;
  // GEMM tile autotuner: chosen configuration per (kind, M, N, K)
  bool autotune = false;
  int force_cfg[2] = {-1, -1};  // 0: projection, 1: half line strength
  int force_split = 1;          // projection split-K when not tuned
  int force_i8_cfg = -1, force_i8_split = 0;  // split-INT8 GEMMs
  int force_crt_cfg = -1;                     // CRT-form int8 GEMMs
  bool crt_kskip = true;  // CRT dipole products skip zero k blocks
  bool crt_fuse = true;   // CRT half ls: reconstruction fused with gather
  bool crt_xfuse = true;  // CRT half ls: X scattered straight into residues
  int proj_crt = 0;  // CRT projection moduli (0: the precision's)
  int f32_cfg = -1;  // FP32 GEMM tile configuration (-1: default)
  std::map<std::tuple<int, int, int, int>, int> tuned;
  std::string tune_log;
  // what the most recent launch of each tuned kernel used (tuning getter)
  std::map<std::string, int64_t> last;
  bool timing = false;
  std::map<std::string, TimerSlot> timers;
  // Bumped whenever device memory a captured sweep may hold is released or
  // changed in place (dipole, replaced block, rebuilt slot tables); a sweep
  // records it at creation and refuses to replay once it differs.
  uint64_t gen = 0;
  // cap (bytes) of the fixed scratch array the allocation-free entry point
  // may use; 0 = the whole array
  int64_t noalloc_cap = 0;
};

namespace gain {

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    GAIN_HIP(hipGetDevice(&prev));
    if (prev != dev) GAIN_HIP(hipSetDevice(dev));
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// RAII event pair around the launches of one kernel family.
struct Timed {
  gain_ctx* c;
  TimerSlot* slot = nullptr;
  std::pair<hipEvent_t, hipEvent_t> ev{};
  Timed(gain_ctx* ctx, const char* name) : c(ctx) {
    if (!c->timing) return;
    slot = &c->timers[name];
    if (!slot->pool.empty()) {
      ev = slot->pool.back();
      slot->pool.pop_back();
    } else {
      GAIN_HIP(hipEventCreate(&ev.first));
      GAIN_HIP(hipEventCreate(&ev.second));
    }
    GAIN_HIP(hipEventRecord(ev.first, c->stream));
  }
  ~Timed() {
    if (!slot) return;
    (void)hipEventRecord(ev.second, c->stream);
    slot->pending.push_back(ev);
  }
};

// Host wall clock of an autotuning pass, added to the "autotune_us"
// tuning counter.
struct TuneClock {
  gain_ctx* c;
  std::chrono::steady_clock::time_point t0;
  explicit TuneClock(gain_ctx* ctx)
      : c(ctx), t0(std::chrono::steady_clock::now()) {}
  ~TuneClock() {
    c->last["autotune_us"] +=
        std::chrono::duration_cast<std::chrono::microseconds>(
            std::chrono::steady_clock::now() - t0)
            .count();
  }
};

struct Plan {
  int jindI, gI, jindF, gF;
  Basis* bI;
  Basis* bF;
  Block* blkI;
  Block* blkF;
  int nI;
  std::vector<int> levels;
  std::vector<double> eI;
  std::vector<uint8_t> okI;
  std::vector<int> DI;                   // initial components computed
  std::vector<std::pair<int, int>> vec;  // (dI, dF) in Output.cpp order
  bool block_ok = false;
  int Flo = 0, Fhi = 0;
};

// Device scratch laid out from one buffer of small per-call arrays.
struct Small {
  int64_t* c_off;
  int* ideg;
  int* h_of_vec;
  int* ideg_of_vec;
  int* slot_of_vec;
  double* eI;
  int* levelI;
  uint8_t* okI;
  int* vptr;  // per H vector: the projection vectors it feeds (CSR)
  int* vidx;
  int* vdeg;
};

// Per-call arrays of one plan, staged to the device.
struct Staged {
  Small sm{};
  const double* cbase = nullptr;
  int nDI = 0, nvpi = 0, nvecX = 0, nvecG = 0;
  bool compute = false;
};

// Half line strengths of every (initial, dI) vector: scatter, mu GEMM and
// gather.  H rows: i * |DI| + k.  Returns the H stride.
gain_ctx* need(gain_ctx* c);
// The whole dipole is resident (every piece uploaded), else GAIN_ERR_STATE.
void need_mu(const gain_ctx* c);
Basis& basis_of(gain_ctx* c, int jind, bool need_infl);
Block& block_of(gain_ctx* c, int jind, int gamma);
void ensure_slots(gain_ctx* c, Basis& b);
// Builds the scatter tables of every uploaded basis whose vibrational
// indices fit the dipole, so later calls need no device allocation.
void build_slots_all(gain_ctx* c);
int mu_ksplit_auto(const gain_ctx* c, int ncolTot);
ScatterArgs make_scatter_args(const gain_ctx* c, const Plan& P,
                              const double* cbase, const Small& sm,
                              int nvecX, double* X);
GemmArgs make_mu_gemm_args(const gain_ctx* c, const double* X, int ncolTot,
                           double* Y, int ks);
GatherArgs make_gather_args(const gain_ctx* c, const Plan& P, int nvecX,
                            int ks, const double* Y, double* H);
// Symmetric-top gather of a plan whose bases are ROTSYM bases (final slot
// tables built, Wigner table uploaded).
GatherRotArgs make_gather_rot_args(gain_ctx* c, const Plan& P, int nvecX,
                                   int ks, const double* Y, double* H,
                                   int64_t ys_alpha, int64_t ys_ks);
// The gather of either kind for a plan (Y in the layout of one mu X GEMM
// over nvecX vectors).
void launch_gather_for(gain_ctx* c, const Plan& P, int nvecX, int ks,
                       const double* Y, double* H, hipStream_t s);
ExpandTArgs make_expand_T_args(const Plan& P, const double* H,
                               const Small& sm, int nvecG, double* G,
                               int64_t gstride);
LinesArgs make_lines_args(const gain_ctx* c, const Plan& P, const Small& sm,
                          int nvpi, bool compute);
// Tile configuration already fixed for a GEMM shape (forced or tuned), -1
// for the default; never times anything.
int gemm_cfg_lookup(const gain_ctx* c, int kind, const GemmArgs& g);
int slices_of(int precision);
int i8_slices_of(int precision);
int crt_moduli_of(int precision);
int crt_cfg_of(gain_ctx* c, int kind, I8GemmArgs g);
int i8_cfg_split(gain_ctx* c, int kind, I8GemmArgs g, int max_split,
                 int* ksplit);
bool degfilter(const gain_ctx* c, int gI, int gF, int dI, int dF);
bool filter_lower(const gain_ctx* c, double e, int g);
bool block_pair_ok(const gain_ctx* c, int jI, int gI, int jF, int gF);
Plan make_plan(gain_ctx* c, int jindI, int gI, const int* levels, int nI,
               int jindF, int gF);
int gemm_cfg(gain_ctx* c, int kind, GemmArgs g);
int gemm_cfg_split(gain_ctx* c, int kind, GemmArgs g, int* ksplit,
                   bool tune_split);
void resolve_timers(gain_ctx* c, TimerSlot& t);
// Resolves this GPU's copy of the allocation-free entry point's scratch
// array (noalloc.hip); called at context creation with `dev` current.
void noalloc_init_device(int dev);

}  // namespace gain
