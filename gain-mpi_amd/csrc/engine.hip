// gain_engine context and C-ABI entry points (the device-manager role of
// GAIN-MPI's GpuManager, BaseClasses/GpuManager.cpp).
//
// One context = one GPU + one HIP stream.  Uploaded data (dipole, bases,
// inflation, eigenvector blocks) stay resident in HBM; per-call work buffers
// grow on demand and are reused.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "engine_internal.hpp"

namespace gain {

static thread_local std::string t_last_error;
void set_last_error(const char* msg) { t_last_error = msg; }

// ---------------------------------------------------------------------------
// Wigner 3j symbol (j1 j2 j3; m1 m2 m3) by the Racah formula in long double,
// the quantity the reference obtains from TROVE's three_j
// (fort_func.f90:7-104) for its table (GpuManager.cpp:164-184).
static long double lfact(int n) {
  long double f = 1.0L;
  for (int i = 2; i <= n; ++i) f *= (long double)i;
  return f;
}

double three_j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (j3 > j1 + j2 || j3 < std::abs(j1 - j2)) return 0.0;
  if (j1 < 0 || j2 < 0 || j3 < 0) return 0.0;
  if (j1 < std::abs(m1) || j2 < std::abs(m2) || j3 < std::abs(m3)) return 0.0;
  if (m1 + m2 + m3 != 0) return 0.0;
  const long double tri = lfact(j1 + j2 - j3) * lfact(j1 - j2 + j3) *
                          lfact(-j1 + j2 + j3) / lfact(j1 + j2 + j3 + 1);
  const long double pre =
      std::sqrt(tri * lfact(j1 + m1) * lfact(j1 - m1) * lfact(j2 + m2) *
                lfact(j2 - m2) * lfact(j3 + m3) * lfact(j3 - m3));
  const int kmin = std::max({0, j2 - j3 - m1, j1 - j3 + m2});
  const int kmax = std::min({j1 + j2 - j3, j1 - m1, j2 + m2});
  long double sum = 0.0L;
  for (int k = kmin; k <= kmax; ++k) {
    const long double den = lfact(k) * lfact(j3 - j2 + k + m1) *
                            lfact(j3 - j1 + k - m2) * lfact(j1 + j2 - j3 - k) *
                            lfact(j1 - k - m1) * lfact(j2 - k + m2);
    sum += (k & 1) ? -1.0L / den : 1.0L / den;
  }
  const int ph = ((j1 - j2 - m3) % 2 + 2) % 2;
  return (double)((ph ? -1.0L : 1.0L) * pre * sum);
}

void resolve_timers(gain_ctx* c, TimerSlot& t) {
  for (auto& e : t.pending) {
    GAIN_HIP(hipEventSynchronize(e.second));
    float ms = 0.f;
    GAIN_HIP(hipEventElapsedTime(&ms, e.first, e.second));
    t.total_ms += ms;
    t.launches += 1;
    t.pool.push_back(e);
  }
  t.pending.clear();
  (void)c;
}

void need_mu(const gain_ctx* c) {
  if (!c->d_mu) fail(GAIN_ERR_STATE, "dipole not uploaded");
  if (c->mu_missing > 0)
    fail(GAIN_ERR_STATE,
         "dipole pieces cover %d of its %d final vibrational columns "
         "(gain_upload_dipole the rest)",
         c->nc - c->mu_missing, c->nc);
}

gain_ctx* need(gain_ctx* c) {
  if (!c) fail(GAIN_ERR_ARG, "null context");
  return c;
}

Basis& basis_of(gain_ctx* c, int jind, bool need_infl) {
  if (jind < 0 || jind >= (int)c->bases.size() || !c->bases[jind].present)
    fail(GAIN_ERR_STATE, "basis for J index %d not uploaded", jind);
  Basis& b = c->bases[jind];
  if (need_infl && !b.has_infl)
    fail(GAIN_ERR_STATE, "inflation data for J index %d not uploaded", jind);
  return b;
}

Block& block_of(gain_ctx* c, int jind, int gamma) {
  auto it = c->blocks.find({jind, gamma});
  if (it == c->blocks.end())
    fail(GAIN_ERR_STATE, "state block (J index %d, irrep %d) not uploaded",
         jind, gamma);
  return it->second;
}

// Scatter CSR for a basis: slot = col*nc + v -> primitives, col = 2K+tau
// (K-block bases) or level * rmd + component (symmetric-top bases).
void ensure_slots(gain_ctx* c, Basis& b) {
  if (b.slot_nc == c->nc) return;
  if (b.nvib > c->nc)
    fail(GAIN_ERR_ARG,
         "vibrational index %d of J=%d exceeds the dipole dimension %d",
         b.nvib - 1, b.J, c->nc);
  const int nslot = b.ncol * c->nc;
  std::vector<int> cnt(nslot + 1, 0);
  std::vector<int> slot(b.nprim);
  for (int r = 0; r < b.nprim; ++r) {
    slot[r] = b.hcol[r] * c->nc + b.hvib[r];
    cnt[slot[r] + 1]++;
  }
  for (int s = 0; s < nslot; ++s) cnt[s + 1] += cnt[s];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1), prim(b.nprim);
  for (int r = 0; r < b.nprim; ++r) prim[fill[slot[r]]++] = r;
  std::vector<int> one(nslot);
  b.slot_multi = false;
  for (int s = 0; s < nslot; ++s) {
    const int n = cnt[s + 1] - cnt[s];
    one[s] = n == 1 ? prim[cnt[s]] : (n == 0 ? -1 : -2);
    b.slot_multi |= n > 1;
  }
  if (b.slot_ptr || b.slot_prim) ++c->gen;
  if (b.slot_ptr) c->arena.release(b.slot_ptr);
  if (b.slot_prim) c->arena.release(b.slot_prim);
  if (b.slot_one) c->arena.release(b.slot_one);
  b.slot_ptr = c->arena.upload(cnt.data(), cnt.size());
  b.slot_prim = c->arena.upload(prim.data(), prim.size());
  b.slot_one = c->arena.upload(one.data(), one.size());
  b.slot_nc = c->nc;
}

void build_slots_all(gain_ctx* c) {
  if (!c->d_mu) return;
  for (Basis& b : c->bases)
    if (b.present && b.nvib <= c->nc) ensure_slots(c, b);
}

int slices_of(int precision) {
  switch (precision) {
    case GAIN_PREC_BF16X3: return 3;
    case GAIN_PREC_BF16X4: return 4;
    default: return 0;
  }
}

int i8_slices_of(int precision) {
  switch (precision) {
    case GAIN_PREC_I8X5: return 5;
    case GAIN_PREC_I8X6: return 6;
    case GAIN_PREC_I8X7: return 7;
    default: return 0;
  }
}

int crt_moduli_of(int precision) {
  switch (precision) {
    case GAIN_PREC_I8CRT11: return 11;
    case GAIN_PREC_I8CRT12: return 12;
    case GAIN_PREC_I8CRT14: return 14;
    default: return 0;
  }
}

void release_sliced(gain_ctx* c, Sliced& sl) {
  if (sl.planes) c->arena.release(sl.planes);
  if (sl.p8) c->arena.release(sl.p8);
  if (sl.expo) c->arena.release(sl.expo);
  if (sl.kr) c->arena.release(sl.kr);
  sl = Sliced{};
}

// Sliced copy of the dipole along v_I (rows = (alpha, v_F)), built once per
// slice count.
const Sliced& mu_sliced(gain_ctx* c, int S) {
  if (c->mu_sl.S == S && !c->mu_sl.i8 && !c->mu_sl.crt) return c->mu_sl;
  Timed t(c, "prep");  // one-time build of a static operand's planes
  release_sliced(c, c->mu_sl);
  const int nc = c->nc;
  double* muT = c->w_tmp.get<double>((size_t)3 * nc * nc);
  transpose_mu(c->d_mu, nc, muT, c->stream);
  Sliced sl;
  sl.S = S;
  sl.Kp = slice_kpad(nc);
  sl.planes = c->arena.alloc<uint16_t>((size_t)S * 3 * nc * sl.Kp);
  sl.expo = c->arena.alloc<int>((size_t)3 * nc * slice_nchunk(S, nc));
  slice_rows_bf16(S, muT, nc, 3 * nc, nc, sl.planes, sl.expo, c->stream);
  c->mu_sl = sl;
  return c->mu_sl;
}

// Sliced copy of a block's eigenvector records, built once per slice count.
const Sliced& block_sliced(gain_ctx* c, Block& b, int S) {
  if (b.sl.S == S && !b.sl.i8 && !b.sl.crt) return b.sl;
  Timed t(c, "prep");  // one-time build of a static operand's planes
  release_sliced(c, b.sl);
  Sliced sl;
  sl.S = S;
  sl.Kp = slice_kpad(b.rec_len);
  sl.planes = c->arena.alloc<uint16_t>((size_t)S * b.n * sl.Kp);
  sl.expo = c->arena.alloc<int>((size_t)b.n * slice_nchunk(S, b.rec_len));
  slice_rows_bf16(S, b.vec, b.ld, b.n, b.rec_len, sl.planes, sl.expo,
                  c->stream);
  b.sl = sl;
  return b.sl;
}

// Split-INT8 copy of the dipole along v_I: one tiled matrix per component
// (rows v_F), built once per slice count.
const Sliced& mu_sliced_i8(gain_ctx* c, int S) {
  if (c->mu_sl.S == S && c->mu_sl.i8 && !c->mu_sl.crt) return c->mu_sl;
  Timed t(c, "prep");  // one-time build of a static operand's planes
  release_sliced(c, c->mu_sl);
  const int nc = c->nc;
  double* muT = c->w_tmp.get<double>((size_t)3 * nc * nc);
  transpose_mu(c->d_mu, nc, muT, c->stream);
  Sliced sl;
  sl.S = S;
  sl.i8 = true;
  sl.Kp = i8_kpad(nc);
  sl.Rp = i8_rows_pad(nc);
  sl.plane = 3 * sl.Rp * sl.Kp;
  sl.p8 = c->arena.alloc<int8_t>((size_t)S * sl.plane);
  sl.expo = c->arena.alloc<int>((size_t)3 * sl.Rp);
  for (int a = 0; a < 3; ++a)
    slice_rows_i8(S, muT + (size_t)a * nc * nc, nc, nc, nc,
                  sl.p8 + a * sl.Rp * sl.Kp, sl.plane, sl.expo + a * sl.Rp,
                  c->stream);
  c->mu_sl = sl;
  return c->mu_sl;
}

// Split-INT8 copy of a block's eigenvector records, built once per slice
// count.
const Sliced& block_sliced_i8(gain_ctx* c, Block& b, int S) {
  if (b.sl.S == S && b.sl.i8 && !b.sl.crt) return b.sl;
  Timed t(c, "prep");  // one-time build of a static operand's planes
  release_sliced(c, b.sl);
  Sliced sl;
  sl.S = S;
  sl.i8 = true;
  sl.Kp = i8_kpad(b.rec_len);
  sl.Rp = i8_rows_pad(b.n);
  sl.plane = sl.Rp * sl.Kp;
  sl.p8 = c->arena.alloc<int8_t>((size_t)S * sl.plane);
  sl.expo = c->arena.alloc<int>((size_t)sl.Rp);
  slice_rows_i8(S, b.vec, b.ld, b.n, b.rec_len, sl.p8, sl.plane, sl.expo,
                c->stream);
  b.sl = sl;
  return b.sl;
}

// Int8 residues of the dipole along v_I (CRT form), one tiled matrix per
// component, built once per moduli count.
const Sliced& mu_sliced_crt(gain_ctx* c, int nm) {
  if (c->mu_sl.S == nm && c->mu_sl.crt) return c->mu_sl;
  Timed t(c, "prep");  // one-time build of a static operand's planes
  release_sliced(c, c->mu_sl);
  const int nc = c->nc;
  double* muT = c->w_tmp.get<double>((size_t)3 * nc * nc);
  transpose_mu(c->d_mu, nc, muT, c->stream);
  Sliced sl;
  sl.S = nm;
  sl.i8 = true;
  sl.crt = true;
  sl.Kp = i8_kpad(nc);
  sl.beta = crt_beta(nm, sl.Kp);
  sl.Rp = i8_rows_pad(nc);
  sl.plane = 3 * sl.Rp * sl.Kp;
  sl.p8 = c->arena.alloc<int8_t>((size_t)nm * sl.plane);
  sl.expo = c->arena.alloc<int>((size_t)3 * sl.Rp);
  for (int a = 0; a < 3; ++a)
    slice_rows_crt(nm, sl.beta, muT + (size_t)a * nc * nc, nc, nc, nc,
                   sl.p8 + a * sl.Rp * sl.Kp, sl.plane, sl.expo + a * sl.Rp,
                   c->stream);
  // the dipole's zero k blocks (entries that round to 0 at beta bits, e.g.
  // far from the diagonal of a banded dipole) are skipped by the products
  const int nrb = (int)(sl.Rp / 32);
  sl.kr = c->arena.alloc<int2>((size_t)3 * nrb);
  i8_block_kranges(sl.p8, sl.plane, nm, sl.Rp * sl.Kp, 3, nrb,
                   (int)(sl.Kp / 32), sl.kr, c->stream);
  sl.kr_host.resize((size_t)3 * nrb);
  GAIN_HIP(hipMemcpyAsync(sl.kr_host.data(), sl.kr,
                        sl.kr_host.size() * sizeof(int2),
                        hipMemcpyDeviceToHost, c->stream));
  GAIN_HIP(hipStreamSynchronize(c->stream));
  c->mu_sl = sl;
  return c->mu_sl;
}

// Int8 residues of a block's eigenvector records (CRT form), built once per
// moduli count.
const Sliced& block_sliced_crt(gain_ctx* c, Block& b, int nm) {
  if (b.sl.S == nm && b.sl.crt) return b.sl;
  Timed t(c, "prep");  // one-time build of a static operand's planes
  release_sliced(c, b.sl);
  Sliced sl;
  sl.S = nm;
  sl.i8 = true;
  sl.crt = true;
  sl.Kp = i8_kpad(b.rec_len);
  sl.beta = crt_beta(nm, sl.Kp);
  sl.Rp = i8_rows_pad(b.n);
  sl.plane = sl.Rp * sl.Kp;
  sl.p8 = c->arena.alloc<int8_t>((size_t)nm * sl.plane);
  sl.expo = c->arena.alloc<int>((size_t)sl.Rp);
  slice_rows_crt(nm, sl.beta, b.vec, b.ld, b.n, b.rec_len, sl.p8, sl.plane,
                 sl.expo, c->stream);
  b.sl = sl;
  return b.sl;
}

// Residue buffer of a CRT-form GEMM, large enough for every configuration.
int8_t* crt_residues(gain_ctx* c, I8GemmArgs g) {
  int64_t mx = 0;
  for (int cfg = 0; cfg < kCrtConfigs; ++cfg) {
    g.cfg = cfg;
    mx = std::max(mx, gemm_crt_res_bytes(g));
  }
  return c->w_res.get<int8_t>((size_t)mx);
}

// Tile configuration of a CRT-form GEMM (kind 0: projection, 1: half line
// strength), timed on the real operands once per shape when autotuning;
// without tuning kCrtDefaultConfig.
int crt_cfg_of(gain_ctx* c, int kind, I8GemmArgs g) {
  if (c->force_crt_cfg >= 0) return c->force_crt_cfg;
  const auto key = std::make_tuple(200 + 10 * kind + g.S,
                                   g.M >= 8192 ? -1 : g.M, g.N, (int)g.Kp);
  auto it = c->tuned.find(key);
  if (it != c->tuned.end()) return it->second;
  if (!c->autotune) return kCrtDefaultConfig;
  TuneClock clk(c);
  hipEvent_t e0, e1;
  GAIN_HIP(hipEventCreate(&e0));
  GAIN_HIP(hipEventCreate(&e1));
  int best = 0;
  float best_ms = 1e30f;
  char line[160];
  snprintf(line, sizeof(line), "crt%d %s M=%d N=%d K=%lld:", g.S,
           kind == 0 ? "projection" : "half_ls", g.M, g.N, (long long)g.Kp);
  std::string rec = line;
  // the residue products, and their reconstruction into FP64 when the
  // caller has an output for it (the fused half-line-strength gather reads
  // the residues itself and passes none)
  auto run = [&] {
    gemm_crt_res(g, c->stream);
    if (g.C) gemm_crt_combine(g, c->stream);
  };
  for (int cfg = 0; cfg < kCrtConfigs; ++cfg) {
    g.cfg = cfg;
    run();  // warm
    GAIN_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < 2; ++r) run();
    GAIN_HIP(hipEventRecord(e1, c->stream));
    GAIN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GAIN_HIP(hipEventElapsedTime(&ms, e0, e1));
    ms *= 0.5f;
    snprintf(line, sizeof(line), " %s=%.3fms", crt_config_name(cfg), ms);
    rec += line;
    if (ms < best_ms) {
      best_ms = ms;
      best = cfg;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  snprintf(line, sizeof(line), " -> %s\n", crt_config_name(best));
  c->tune_log += rec + line;
  c->tuned[key] = best;
  return best;
}

// Tile configuration and split count of a split-INT8 GEMM (kind 0:
// projection, 1: half line strength), timed on the real operands once per
// shape when autotuning (splits from the smallest exact one up to
// `max_split`; g.C must hold that many partials).  Without tuning: the
// 64 x 128 tile and the smallest exact split.
int i8_cfg_split(gain_ctx* c, int kind, I8GemmArgs g, int max_split,
                 int* ksplit) {
  const int64_t nkb = g.Kp / 32, maxkb = i8_max_krange(g.S) / 32;
  const int min_split = (int)((nkb + maxkb - 1) / maxkb);
  GAIN_CHECK_ARG(max_split >= min_split,
                 "int8 gemm: K = %lld needs %d splits", (long long)g.Kp,
                 min_split);
  if (c->force_i8_cfg >= 0) {
    GAIN_CHECK_ARG(i8_config_ok(c->force_i8_cfg, g.S),
                   "int8 configuration %d does not run %d slices",
                   c->force_i8_cfg, g.S);
    *ksplit = std::max(min_split, std::min(max_split, c->force_i8_split));
    return c->force_i8_cfg;
  }
  const auto key = std::make_tuple(100 + 10 * kind + g.S,
                                   g.M >= 8192 ? -1 : g.M, g.N, (int)g.Kp);
  auto it = c->tuned.find(key);
  if (it != c->tuned.end()) {
    *ksplit = it->second / 100;
    return it->second % 100;
  }
  *ksplit = min_split;
  if (!c->autotune) return 0;
  TuneClock clk(c);
  hipEvent_t e0, e1;
  GAIN_HIP(hipEventCreate(&e0));
  GAIN_HIP(hipEventCreate(&e1));
  int best = 0, best_ks = min_split;
  float best_ms = 1e30f;
  char line[160];
  snprintf(line, sizeof(line), "int8x%d %s M=%d N=%d K=%lld:", g.S,
           kind == 0 ? "projection" : "half_ls", g.M, g.N, (long long)g.Kp);
  std::string rec = line;
  for (int cfg = 0; cfg < kI8Configs; ++cfg)
    for (int ks = min_split; ks <= max_split && i8_config_ok(cfg, g.S);
         ks *= 2) {
      g.cfg = cfg;
      g.ksplit = ks;
      gemm_sliced_i8(g, c->stream);  // warm
      GAIN_HIP(hipEventRecord(e0, c->stream));
      for (int r = 0; r < 2; ++r) gemm_sliced_i8(g, c->stream);
      GAIN_HIP(hipEventRecord(e1, c->stream));
      GAIN_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      GAIN_HIP(hipEventElapsedTime(&ms, e0, e1));
      ms *= 0.5f;
      snprintf(line, sizeof(line), " %s/k%d=%.3fms", i8_config_name(cfg), ks,
               ms);
      rec += line;
      if (ms < best_ms) {
        best_ms = ms;
        best = cfg;
        best_ks = ks;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  snprintf(line, sizeof(line), " -> %s split %d\n", i8_config_name(best),
           best_ks);
  c->tune_log += rec + line;
  c->tuned[key] = best_ks * 100 + best;
  *ksplit = best_ks;
  return best;
}

bool degfilter(const gain_ctx* c, int gI, int gF, int dI, int dF) {
  // States::DegeneracyFilter, States.cpp:95-116
  const gain_params& p = c->params;
  if (p.sym_degen[gI] == 1 || p.sym_degen[gF] == 1) return true;
  if (!p.reduced) return true;
  if (p.symmetry_type == 1) return true;
  if (p.symmetry_type == 2) return dI == 1 && dF == 0;
  if (p.symmetry_type == 3) {
    if (gI == 3 && gF == 4 && dI == 1 && dF == 0) return true;
    if (gI == 4 && gF == 3 && dI == 2 && dF == 0) return true;
    if (gI == 2 && gF == 2 && dI == 1 && dF == 0) return true;
  }
  return false;
}

// Block-level selection rules of a (J_I, gamma_I) -> (J_F, gamma_F) pair:
// gns > 0 on both sides, not both J = maxJ, |dJ| <= 1, J_I + J_F >= 1
// (States.cpp:54-76) and the irrep pair (TroveInput.cpp:202-236).
bool block_pair_ok(const gain_ctx* c, int jI, int gI, int jF, int gF) {
  const gain_params& p = c->params;
  return p.gns[gI] > 0.0 && p.gns[gF] > 0.0 &&
         (jF != p.maxJ || jI != p.maxJ) &&
         p.isym_pairs[gI] == p.isym_pairs[gF] && c->pair[gI] == gF &&
         std::abs(jF - jI) <= 1 && jI + jF >= 1;
}

bool filter_lower(const gain_ctx* c, double e, int g) {
  const gain_params& p = c->params;
  return (e - p.zpe) <= p.erange_lower[1] && (e - p.zpe) >= p.erange_lower[0] &&
         p.isym_do[g];
}

// First index in [0,n) where pred holds, for a monotone predicate.
template <class P>
int lower_bound_pred(int n, P pred) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (pred(mid))
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Rows of the final block that can pass the energy part of the filter for
// any of the initial energies (superset; the exact test runs per pair).
void final_range(const gain_ctx* c, const Block& bF,
                 const std::vector<double>& eI, const std::vector<uint8_t>& ok,
                 int* lo, int* hi) {
  const gain_params& p = c->params;
  if (!bF.sorted) {
    *lo = 0;
    *hi = bF.n;
    return;
  }
  int L = bF.n, H = 0;
  const std::vector<double>& e = bF.energy;
  for (size_t i = 0; i < eI.size(); ++i) {
    if (!ok[i]) continue;
    const double ei = eI[i];
    // nu >= fw0 and eF - zpe >= eu0 are monotone in eF
    const int l = lower_bound_pred(bF.n, [&](int f) {
      return (e[f] - ei) >= p.freq_window[0] && (e[f] - p.zpe) >= p.erange_upper[0];
    });
    // first index where nu > fw1 or eF - zpe > eu1
    const int h = lower_bound_pred(bF.n, [&](int f) {
      return !((e[f] - ei) <= p.freq_window[1]) ||
             !((e[f] - p.zpe) <= p.erange_upper[1]);
    });
    if (l < h) {
      L = std::min(L, l);
      H = std::max(H, h);
    }
  }
  if (L >= H) L = H = 0;
  *lo = L;
  *hi = H;
}

Plan make_plan(gain_ctx* c, int jindI, int gI, const int* levels, int nI,
               int jindF, int gF) {
  if (!c->params_set) fail(GAIN_ERR_STATE, "gain_set_params not called");
  need_mu(c);
  GAIN_CHECK_ARG(gI >= 0 && gI < c->nrepres && gF >= 0 && gF < c->nrepres,
                 "irrep index out of range");
  Plan P;
  P.jindI = jindI;
  P.gI = gI;
  P.jindF = jindF;
  P.gF = gF;
  P.bI = &basis_of(c, jindI, true);
  P.bF = &basis_of(c, jindF, true);
  if (P.bI->rotsym != P.bF->rotsym)
    fail(GAIN_ERR_ARG, "J indices %d and %d mix a symmetric-top and a "
         "K-block basis", jindI, jindF);
  P.blkI = &block_of(c, jindI, gI);
  P.blkF = &block_of(c, jindF, gF);
  if (!P.blkF->vec)
    fail(GAIN_ERR_STATE, "final block (%d,%d) has no eigenvectors", jindF, gF);
  GAIN_CHECK_ARG(P.blkI->rec_len == P.bI->irrep[gI].ntot,
                 "initial block record length %d != Ntot %d",
                 P.blkI->rec_len, P.bI->irrep[gI].ntot);
  GAIN_CHECK_ARG(P.blkF->rec_len == P.bF->irrep[gF].ntot,
                 "final block record length %d != Ntot %d", P.blkF->rec_len,
                 P.bF->irrep[gF].ntot);
  GAIN_CHECK_ARG(P.bI->J <= c->jmax && P.bF->J <= c->jmax,
                 "J above the context's jmax %d", c->jmax);
  GAIN_CHECK_ARG(c->params.sym_degen[gI] == P.bI->irrep[gI].sdeg &&
                     c->params.sym_degen[gF] == P.bF->irrep[gF].sdeg,
                 "irrep degeneracy of the parameters differs from the "
                 "inflation data");
  GAIN_CHECK_ARG(nI >= 1, "no initial states");
  P.nI = nI;
  const gain_params& p = c->params;
  for (int i = 0; i < nI; ++i) {
    GAIN_CHECK_ARG(levels[i] >= 0 && levels[i] < P.blkI->n,
                   "initial level %d out of range [0,%d)", levels[i],
                   P.blkI->n);
    P.levels.push_back(levels[i]);
    const double e = P.blkI->energy[levels[i]];
    P.eI.push_back(e);
    P.okI.push_back(filter_lower(c, e, gI) ? 1 : 0);
  }
  P.block_ok = block_pair_ok(c, P.bI->J, gI, P.bF->J, gF);
  const int ndegI = p.sym_degen[gI], ndegF = p.sym_degen[gF];
  for (int dI = 0; dI < ndegI; ++dI)
    if (degfilter(c, gI, c->pair[gI], dI, 0)) P.DI.push_back(dI);
  for (int dF = 0; dF < ndegF; ++dF)
    for (int dI = 0; dI < ndegI; ++dI)
      if (degfilter(c, gI, gF, dI, dF)) P.vec.push_back({dI, dF});
  if (P.block_ok) final_range(c, *P.blkF, P.eI, P.okI, &P.Flo, &P.Fhi);
  return P;
}

// Tile configuration for a GEMM of one kind and shape: forced, tuned (the
// autotuner times every configuration on the real operands once per shape,
// on the context's stream), or the default.
int gemm_cfg_lookup(const gain_ctx* c, int kind, const GemmArgs& g) {
  if (c->force_cfg[kind] >= 0) return c->force_cfg[kind];
  const auto key = std::make_tuple(kind, g.M >= 8192 ? -1 : g.M, g.N, g.K);
  auto it = c->tuned.find(key);
  return it != c->tuned.end() ? it->second % 100 : -1;
}

int gemm_cfg(gain_ctx* c, int kind, GemmArgs g) {
  int ks = 0;
  return gemm_cfg_split(c, kind, g, &ks, false);
}

// Same, and for kind 0 (projection) also the split-K count when
// `tune_split` (the caller then sums the partials).
int gemm_cfg_split(gain_ctx* c, int kind, GemmArgs g, int* ksplit,
                   bool tune_split) {
  if (c->force_cfg[kind] >= 0) {
    if (tune_split) *ksplit = std::max(1, c->force_split);
    return c->force_cfg[kind];
  }
  // large M only changes the number of row tiles: one choice per N, K
  const auto key = std::make_tuple(kind, g.M >= 8192 ? -1 : g.M, g.N, g.K);
  auto it = c->tuned.find(key);
  if (it != c->tuned.end()) {
    if (tune_split) *ksplit = it->second / 100;
    return it->second % 100;
  }
  if (!c->autotune) {
    if (tune_split) *ksplit = std::max(1, c->force_split);
    return -1;
  }
  TuneClock clk(c);
  hipEvent_t e0, e1;
  GAIN_HIP(hipEventCreate(&e0));
  GAIN_HIP(hipEventCreate(&e1));
  int best = -1, best_ks = 1;
  float best_ms = 1e30f;
  char line[160];
  snprintf(line, sizeof(line), "%s M=%d N=%d K=%d:",
           kind == 0 ? "projection" : "half_ls", g.M, g.N, g.K);
  std::string rec = line;
  const int splits[4] = {1, 2, 4, 8};
  const int nsplit = tune_split ? 4 : 1;
  for (int cfg = 0; cfg < kGemmConfigs; ++cfg)
    for (int si = 0; si < nsplit; ++si) {
      g.cfg = cfg;
      if (tune_split) {
        g.ksplit = splits[si];
        g.sCks = (int64_t)g.ldc * g.N;
      }
      gemm_f64(g, c->stream);  // warm
      GAIN_HIP(hipEventRecord(e0, c->stream));
      for (int r = 0; r < 2; ++r) gemm_f64(g, c->stream);
      GAIN_HIP(hipEventRecord(e1, c->stream));
      GAIN_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      GAIN_HIP(hipEventElapsedTime(&ms, e0, e1));
      ms *= 0.5f;
      if (tune_split)
        snprintf(line, sizeof(line), " %s/k%d=%.3fms", gemm_config_name(cfg),
                 g.ksplit, ms);
      else
        snprintf(line, sizeof(line), " %s=%.3fms", gemm_config_name(cfg), ms);
      rec += line;
      if (ms < best_ms) {
        best_ms = ms;
        best = cfg;
        best_ks = tune_split ? g.ksplit : 1;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  snprintf(line, sizeof(line), " -> %s split %d\n", gemm_config_name(best),
           best_ks);
  c->tune_log += rec + line;
  c->tuned[key] = best_ks * 100 + best;
  if (tune_split) *ksplit = best_ks;
  return best;
}

// Split-K count of the FP64 mu X GEMM when none is forced: enough
// workgroups to fill 256 CUs (64-row tiles x column tiles x 3 components).
int mu_ksplit_auto(const gain_ctx* c, int ncolTot) {
  if (c->mu_ksplit > 0) return c->mu_ksplit;
  const int nc = c->nc;
  const int64_t tiles = (int64_t)((nc + 63) / 64) *
                        ((ncolTot + (ncolTot > 64 ? 127 : 63)) /
                         (ncolTot > 64 ? 128 : 64)) * 3;
  return (int)std::min<int64_t>(16, std::max<int64_t>(1, 2048 / tiles));
}

// Scatter + expansion of nvecX initial vectors into X[vec][2K+tau][v].
ScatterArgs make_scatter_args(const gain_ctx* c, const Plan& P,
                              const double* cbase, const Small& sm,
                              int nvecX, double* X) {
  const Basis& bI = *P.bI;
  const Irrep& ir = bI.irrep[P.gI];
  ScatterArgs sa{};
  sa.nslot = bI.ncol * c->nc;
  sa.nprim = bI.nprim;
  sa.slot_ptr = bI.slot_ptr;
  sa.slot_prim = bI.slot_prim;
  sa.kval = bI.kval;
  sa.tau = bI.tau;
  sa.icontr = bI.icontr;
  sa.ijt = bI.ijterms + (int64_t)P.gI * bI.maxsymcoeffs;
  sa.N = ir.N;
  sa.repres = ir.repres;
  sa.sdeg = ir.sdeg;
  sa.ntot = ir.ntot;
  sa.cbase = cbase;
  sa.c_off = sm.c_off;
  sa.ideg = sm.ideg;
  sa.nvec = nvecX;
  sa.X = X;
  sa.xstride = (int64_t)bI.ncol * c->nc;
  return sa;
}

// FP64 mu X GEMM: Y[ks][alpha][col][vF] for ncolTot columns of X.
GemmArgs make_mu_gemm_args(const gain_ctx* c, const double* X, int ncolTot,
                           double* Y, int ks) {
  const int nc = c->nc;
  GemmArgs g{};
  g.M = nc;
  g.N = ncolTot;
  g.K = nc;
  g.A = c->d_mu;
  g.lda = nc;
  g.sAz = (int64_t)nc * nc;
  g.a_kcontig = false;
  g.B = X;
  g.ldb = nc;
  g.C = Y;
  g.ldc = nc;
  g.sCz = (int64_t)ncolTot * nc;
  g.sCks = 3 * (int64_t)ncolTot * nc;
  g.nz = 3;
  g.ksplit = ks;
  return g;
}

// Gather of the half line strengths H[vec][r'] from Y (3j, rotational
// factor and signs per final primitive).
GatherArgs make_gather_args(const gain_ctx* c, const Plan& P, int nvecX,
                            int ks, const double* Y, double* H) {
  const Basis& bI = *P.bI;
  const Basis& bF = *P.bF;
  const int64_t ncolTot = (int64_t)nvecX * bI.ncol;
  GatherArgs ga{};
  ga.nprimF = bF.nprim;
  ga.nc = c->nc;
  ga.jI = bI.J;
  ga.jF = bF.J;
  ga.jmax = c->jmax;
  ga.ncolI = bI.ncol;
  ga.nvec = nvecX;
  ga.ksplit = ks;
  ga.kvalF = bF.kval;
  ga.tauF = bF.tau;
  ga.vibF = bF.vib;
  ga.threej = c->d_threej;
  ga.Y = Y;
  ga.ys_alpha = ncolTot * c->nc;
  ga.ys_ks = 3 * ncolTot * c->nc;
  ga.H = H;
  ga.hstride = bF.nprim;
  return ga;
}

GatherRotArgs make_gather_rot_args(gain_ctx* c, const Plan& P, int nvecX,
                                   int ks, const double* Y, double* H,
                                   int64_t ys_alpha, int64_t ys_ks) {
  Basis& bI = *P.bI;
  Basis& bF = *P.bF;
  auto it = c->wigner.find({P.jindI, P.jindF});
  if (it == c->wigner.end())
    fail(GAIN_ERR_STATE,
         "no rotational (Wigner) table for the symmetric-top pair of J "
         "indices %d -> %d (gain_upload_wigner)",
         P.jindI, P.jindF);
  const WignerTable& w = it->second;
  ensure_slots(c, bF);
  GatherRotArgs ga{};
  ga.nc = c->nc;
  ga.ncolI = bI.ncol;
  ga.ncolF = bF.ncol;
  ga.nvec = nvecX;
  ga.ksplit = ks;
  ga.nlI = w.nlI;
  ga.nlF = w.nlF;
  ga.md = w.md;
  ga.W = w.d;
  ga.Y = Y;
  ga.ys_alpha = ys_alpha;
  ga.ys_ks = ys_ks;
  ga.slot_ptrF = bF.slot_ptr;
  ga.slot_primF = bF.slot_prim;
  ga.H = H;
  ga.hstride = bF.nprim;
  return ga;
}

void launch_gather_for(gain_ctx* c, const Plan& P, int nvecX, int ks,
                       const double* Y, double* H, hipStream_t s) {
  if (P.bI->rotsym) {
    const int64_t ncolTot = (int64_t)nvecX * P.bI->ncol;
    launch_gather_rotsym(make_gather_rot_args(c, P, nvecX, ks, Y, H,
                                              ncolTot * c->nc,
                                              3 * ncolTot * c->nc),
                         s);
  } else {
    launch_gather_h(make_gather_args(c, P, nvecX, ks, Y, H), s);
  }
}

// Transposed expansion G[v] = P_{F,dF(v)}^T H[h_of_vec[v]].
ExpandTArgs make_expand_T_args(const Plan& P, const double* H,
                               const Small& sm, int nvecG, double* G,
                               int64_t gstride) {
  const Basis& bF = *P.bF;
  const Irrep& irF = bF.irrep[P.gF];
  ExpandTArgs ea{};
  ea.ntot = irF.ntot;
  ea.nprim = bF.nprim;
  ea.sdeg = irF.sdeg;
  ea.tcase = irF.tcase;
  ea.case_ptr = bF.case_ptr;
  ea.case_prim = bF.case_prim;
  ea.icontr = bF.icontr;
  ea.repres = irF.repres;
  ea.H = H;
  ea.hstride = bF.nprim;
  ea.h_of_vec = sm.h_of_vec;
  ea.ideg_of_vec = sm.ideg_of_vec;
  ea.nvec = nvecG;
  ea.G = G;
  ea.gstride = gstride;
  return ea;
}

// Epilogue arguments (filters, S, A, intensity) of one plan; the caller
// sets the projected-strength source and the outputs.
LinesArgs make_lines_args(const gain_ctx* c, const Plan& P, const Small& sm,
                          int nvpi, bool compute) {
  const gain_params& p = c->params;
  LinesArgs la{};
  la.nI = P.nI;
  la.nF = P.blkF->n;
  la.Flo = compute ? P.Flo : 0;
  la.Fhi = compute ? P.Fhi : 0;
  la.maxdeg = c->maxdeg;
  la.nvec_per_i = nvpi;
  la.slot_of_vec = sm.slot_of_vec;
  la.ksplit = 1;
  la.eI = sm.eI;
  la.levelI = sm.levelI;
  la.okI = sm.okI;
  la.eF = P.blkF->d_energy;
  la.same_block = (P.jindI == P.jindF && P.gI == P.gF);
  la.block_ok = P.block_ok;
  la.jI = P.bI->J;
  la.jF = P.bF->J;
  la.ndegI = p.sym_degen[P.gI];
  la.ndegF = p.sym_degen[P.gF];
  la.reduced = p.reduced;
  la.gnsF = p.gns[P.gF];
  la.zpe = p.zpe;
  la.threshold = p.threshold;
  la.beta = kPlanck * kVellgt / (kBoltz * p.temperature);
  la.partition = p.partition;
  la.fw0 = p.freq_window[0];
  la.fw1 = p.freq_window[1];
  la.eu0 = p.erange_upper[0];
  la.eu1 = p.erange_upper[1];
  la.isym_do_F = p.isym_do[P.gF];
  return la;
}

int64_t run_half_ls(gain_ctx* c, const Plan& P, const double* cbase,
                    const Small& sm, int nvecX, double* H, int precision) {
  Basis& bI = *P.bI;
  Basis& bF = *P.bF;
  ensure_slots(c, bI);
  if (bF.nvib > c->nc)
    fail(GAIN_ERR_ARG,
         "vibrational index %d of J=%d exceeds the dipole dimension %d",
         bF.nvib - 1, bF.J, c->nc);
  const int nc = c->nc;
  const int64_t xstride = (int64_t)bI.ncol * nc;
  double* X = c->w_X.get<double>((size_t)nvecX * xstride);
  const int ncolTot = nvecX * bI.ncol;
  int ks = mu_ksplit_auto(c, ncolTot);
  // the CRT form sizes its own Y (none when the gather is fused)
  double* Y = crt_moduli_of(precision) > 0
                  ? nullptr
                  : c->w_Y.get<double>((size_t)ks * 3 * ncolTot * nc);
  Timed t(c, "half_ls");
  ScatterArgs sa = make_scatter_args(c, P, cbase, sm, nvecX, X);
  // CRT form: X built straight into its residue planes when it fits
  const int64_t repres_len =
      (int64_t)bI.matsize * bI.irrep[P.gI].sdeg * bI.irrep[P.gI].ntot;
  const bool xfuse = crt_moduli_of(precision) > 0 && c->crt_xfuse &&
                     scatter_slice_crt_fits(crt_moduli_of(precision), nc,
                                            repres_len);
  if (!xfuse) launch_scatter_expand(sa, c->stream);
  const int S = slices_of(precision);
  if (S > 0) {
    ks = 1;
    const Sliced& ms = mu_sliced(c, S);
    uint16_t* xs = c->w_xs.get<uint16_t>((size_t)S * ncolTot * ms.Kp);
    const int nch = (int)slice_nchunk(S, nc);
    int* xe = c->w_xe.get<int>((size_t)ncolTot * nch);
    slice_rows_bf16(S, X, nc, ncolTot, nc, xs, xe, c->stream);
    SlicedGemmArgs g{};
    g.S = S;
    g.M = nc;
    g.N = ncolTot;
    g.Kp = ms.Kp;
    g.A = ms.planes;
    g.a_plane = (int64_t)3 * nc * ms.Kp;
    g.a_z = (int64_t)nc * ms.Kp;
    g.eA = ms.expo;
    g.ea_z = (int64_t)nc * nch;
    g.nch = nch;
    g.B = xs;
    g.b_plane = (int64_t)ncolTot * ms.Kp;
    g.eB = xe;
    g.C = Y;
    g.ldc = nc;
    g.c_z = (int64_t)ncolTot * nc;
    g.nz = 3;
    {
      Timed tg(c, "half_ls_gemm");
      gemm_sliced_bf16(g, c->stream);
    }
  } else if (crt_moduli_of(precision) > 0) {
    const int nm = crt_moduli_of(precision);
    const Sliced& ms = mu_sliced_crt(c, nm);
    const int64_t Rx = i8_rows_pad(ncolTot), plane = Rx * ms.Kp;
    int8_t* xs = c->w_xs.get<int8_t>((size_t)nm * plane);
    int* xe = c->w_xe.get<int>((size_t)Rx);
    if (xfuse)
      scatter_slice_rows_crt(nm, ms.beta, sa, bI.ncol, nc, ncolTot, xs, plane,
                             xe, c->w_desc.get<int4>((size_t)sa.nslot),
                             !bI.slot_multi && bI.irrep[P.gI].maxN <= 1,
                             c->stream);
    else
      slice_rows_crt(nm, ms.beta, X, nc, ncolTot, nc, xs, plane, xe,
                     c->stream);
    I8GemmArgs g{};
    g.S = nm;
    g.beta = ms.beta;
    g.M = nc;
    g.N = ncolTot;
    g.row0 = 0;
    g.Kp = ms.Kp;
    g.A = ms.p8;
    g.a_plane = ms.plane;
    g.a_z = ms.Rp * ms.Kp;
    g.eA = ms.expo;
    g.ea_z = ms.Rp;
    g.B = xs;
    g.b_plane = plane;
    g.eB = xe;
    g.ldc = nc;
    g.c_z = (int64_t)ncolTot * nc;
    g.nz = 3;
    if (c->crt_kskip) {
      g.a_kr = ms.kr;
      g.akr_z = ms.Rp / 32;
      g.akr_n = (int)(ms.Rp / 32);
    }
    ks = 1;
    // K-block bases: reconstruction fused with the gather (no FP64 Y)
    const bool fuse = c->crt_fuse && !bF.rotsym;
    Y = fuse ? nullptr : c->w_Y.get<double>((size_t)3 * ncolTot * nc);
    g.C = Y;
    g.R = crt_residues(c, g);
    g.cfg = crt_cfg_of(c, 1, g);
    // k blocks multiplied, parts per million of the dense count
    c->last["half_ls_kblocks_ppm"] = (int64_t)std::llround(
        1e6 * crt_kblock_fraction(g, ms.kr_host.data()));
    {
      Timed tg(c, "half_ls_gemm");
      gemm_crt_res(g, c->stream);
    }
    c->last["half_ls_cfg"] = g.cfg;
    c->last["half_ls_split"] = 1;
    c->last["half_ls_grid"] = gemm_crt_grid_items(g);
    if (fuse) {
      ensure_slots(c, bF);
      CrtGatherArgs ga{};
      ga.g = g;
      ga.nc = nc;
      ga.nvec = nvecX;
      ga.ncolI = bI.ncol;
      ga.jI = bI.J;
      ga.jF = bF.J;
      ga.jmax = c->jmax;
      ga.threej = c->d_threej;
      ga.slot_ptr = bF.slot_ptr;
      ga.slot_prim = bF.slot_prim;
      ga.slot_one = bF.slot_one;
      ga.H = H;
      ga.hstride = bF.nprim;
      launch_crt_gather_h(ga, c->stream);
      return bF.nprim;
    }
    gemm_crt_combine(g, c->stream);
  } else if (i8_slices_of(precision) > 0) {
    const int S8 = i8_slices_of(precision);
    const Sliced& ms = mu_sliced_i8(c, S8);
    const int64_t Rx = i8_rows_pad(ncolTot), plane = Rx * ms.Kp;
    int8_t* xs = c->w_xs.get<int8_t>((size_t)S8 * plane);
    int* xe = c->w_xe.get<int>((size_t)Rx);
    slice_rows_i8(S8, X, nc, ncolTot, nc, xs, plane, xe, c->stream);
    I8GemmArgs g{};
    g.S = S8;
    g.M = nc;
    g.N = ncolTot;
    g.row0 = 0;
    g.Kp = ms.Kp;
    g.A = ms.p8;
    g.a_plane = ms.plane;
    g.a_z = ms.Rp * ms.Kp;
    g.eA = ms.expo;
    g.ea_z = ms.Rp;
    g.B = xs;
    g.b_plane = plane;
    g.eB = xe;
    g.ldc = nc;
    g.c_z = (int64_t)ncolTot * nc;
    g.c_ks = 3 * (int64_t)ncolTot * nc;
    g.nz = 3;
    // K = Nc is far below the exact range: no split
    Y = c->w_Y.get<double>((size_t)3 * ncolTot * nc);
    g.C = Y;
    g.cfg = i8_cfg_split(c, 1, g, 1, &ks);
    g.ksplit = ks;
    {
      Timed tg(c, "half_ls_gemm");
      gemm_sliced_i8(g, c->stream);
    }
    c->last["half_ls_cfg"] = g.cfg;
    c->last["half_ls_split"] = ks;
    c->last["half_ls_grid"] = gemm_i8_grid_items(g);
  } else if (precision == GAIN_PREC_FP32) {
    ks = 1;
    GemmArgs g{};
    g.M = nc;
    g.N = ncolTot;
    g.K = nc;
    g.A = c->d_mu;
    g.lda = nc;
    g.sAz = (int64_t)nc * nc;
    g.a_kcontig = false;
    g.B = X;
    g.ldb = nc;
    g.C = Y;
    g.ldc = nc;
    g.sCz = (int64_t)ncolTot * nc;
    g.nz = 3;
    g.ksplit = 1;
    {
      Timed tg(c, "half_ls_gemm");
      gemm_f32acc(g, c->stream, c->f32_cfg);
    }
  } else if (c->use_mfma) {
    GemmArgs g = make_mu_gemm_args(c, X, ncolTot, Y, ks);
    g.cfg = gemm_cfg(c, 1, g);
    {
      Timed tg(c, "half_ls_gemm");
      gemm_f64(g, c->stream);
    }
    c->last["half_ls_cfg"] = g.cfg < 0 ? gemm_default_config(g.N) : g.cfg;
    c->last["half_ls_split"] = ks;
    c->last["half_ls_grid"] = gemm_grid_items(g);
  } else {
    launch_mu_gemm_valu(c->d_mu, nc, X, ncolTot, Y, ks, c->stream);
  }
  launch_gather_for(c, P, nvecX, ks, Y, H, c->stream);
  return bF.nprim;
}

Staged stage(gain_ctx* c, const Plan& P, const double* cI_host) {
  Block& blkI = *P.blkI;
  Staged st;
  const int nI = P.nI;
  const int nDI = (int)P.DI.size();
  const int nvpi = (int)P.vec.size();
  const int nvecX = nI * nDI;
  const int nvecG = nI * nvpi;
  st.nDI = nDI;
  st.nvpi = nvpi;
  st.nvecX = nvecX;
  st.nvecG = nvecG;
  st.compute = P.Fhi > P.Flo && nvpi > 0 && nDI > 0;
  // ---- small per-call arrays through the pinned ring
  const size_t off_c = 0;
  const size_t off_ideg = off_c + sizeof(int64_t) * std::max(nvecX, 1);
  const size_t off_hv = off_ideg + sizeof(int) * std::max(nvecX, 1);
  const size_t off_dv = off_hv + sizeof(int) * std::max(nvecG, 1);
  const size_t off_sv = off_dv + sizeof(int) * std::max(nvecG, 1);
  const size_t off_vp = off_sv + sizeof(int) * 16;
  const size_t off_vi = off_vp + sizeof(int) * (nvecX + 2);
  const size_t off_vd = off_vi + sizeof(int) * std::max(nvecG, 1);
  const size_t off_e = (off_vd + sizeof(int) * std::max(nvecG, 1) + 7) / 8 * 8;
  const size_t off_l = off_e + sizeof(double) * nI;
  const size_t off_ok = off_l + sizeof(int) * nI;
  const size_t total = off_ok + nI + 16;
  int slot;
  char* hs = (char*)c->ring.acquire(total, &slot);
  int64_t* h_c = (int64_t*)(hs + off_c);
  int* h_ideg = (int*)(hs + off_ideg);
  int* h_hv = (int*)(hs + off_hv);
  int* h_dv = (int*)(hs + off_dv);
  int* h_sv = (int*)(hs + off_sv);
  int* h_vp = (int*)(hs + off_vp);
  int* h_vi = (int*)(hs + off_vi);
  int* h_vd = (int*)(hs + off_vd);
  for (int i = 0; i < nI; ++i)
    for (int k = 0; k < nDI; ++k) {
      h_c[i * nDI + k] = cI_host ? 0 : (int64_t)P.levels[i] * blkI.ld;
      h_ideg[i * nDI + k] = P.DI[k];
    }
  for (int i = 0; i < nI; ++i)
    for (int v = 0; v < nvpi; ++v) {
      const int dI = P.vec[v].first, dF = P.vec[v].second;
      int pos = -1;
      for (int k = 0; k < nDI; ++k)
        if (P.DI[k] == dI) pos = k;
      // a (dI,dF) pair whose dI has no half line strength reads a zero
      // vector (the reference reads its memset-zero buffer)
      h_hv[i * nvpi + v] = pos >= 0 ? i * nDI + pos : nvecX;
      h_dv[i * nvpi + v] = dF;
    }
  for (int v = 0; v < nvpi; ++v)
    h_sv[v] = P.vec[v].first + P.vec[v].second * c->maxdeg;
  // inverse of h_of_vec (for the adjoint): vectors fed by each H vector
  int q = 0;
  for (int hv = 0; hv < nvecX; ++hv) {
    h_vp[hv] = q;
    for (int v = 0; v < nvecG; ++v)
      if (h_hv[v] == hv) {
        h_vi[q] = v;
        h_vd[q] = h_dv[v];
        ++q;
      }
  }
  h_vp[nvecX] = q;
  std::memcpy(hs + off_e, P.eI.data(), sizeof(double) * nI);
  std::memcpy(hs + off_l, P.levels.data(), sizeof(int) * nI);
  std::memcpy(hs + off_ok, P.okI.data(), nI);
  char* ds = c->w_small.get<char>(total);
  GAIN_HIP(hipMemcpyAsync(ds, hs, total, hipMemcpyHostToDevice, c->stream));
  st.sm = Small{(int64_t*)(ds + off_c), (int*)(ds + off_ideg),
                (int*)(ds + off_hv),    (int*)(ds + off_dv),
                (int*)(ds + off_sv),    (double*)(ds + off_e),
                (int*)(ds + off_l),     (uint8_t*)(ds + off_ok),
                (int*)(ds + off_vp),    (int*)(ds + off_vi),
                (int*)(ds + off_vd)};
  st.cbase = blkI.vec;
  int cslot = -1;
  if (cI_host) {
    const size_t nb = sizeof(double) * blkI.rec_len;
    void* hc = c->ring.acquire(nb, &cslot);
    std::memcpy(hc, cI_host, nb);
    double* dc = c->w_c.get<double>(blkI.rec_len);
    GAIN_HIP(hipMemcpyAsync(dc, hc, nb, hipMemcpyHostToDevice, c->stream));
    st.cbase = dc;
  }
  c->ring.release(slot, c->stream);
  if (cslot >= 0) c->ring.release(cslot, c->stream);
  if (!st.cbase && st.compute)
    fail(GAIN_ERR_STATE, "initial block has no eigenvectors and no cI given");
  return st;
}

// GEMV autotuner: rows per wave x split-K count, timed on the real final
// block once per (row count bucket, record length, vector count).
void gemv_tune(gain_ctx* c, const double* C, int64_t ld, int ntot, int F0,
               int nR, const double* G, int64_t gstride, int nvec) {
  const auto key = std::make_tuple(2, nR >= 8192 ? -1 : nR, nvec, ntot);
  if (c->tuned.count(key)) return;
  TuneClock clk(c);
  const int rows_opt[3] = {2, 4, 8};
  const int ks_opt[5] = {2, 4, 8, 16, 32};
  hipEvent_t e0, e1;
  GAIN_HIP(hipEventCreate(&e0));
  GAIN_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int br = c->gemv_rows, bk = c->gemv_ksplit;
  char line[96];
  snprintf(line, sizeof(line), "gemv rows=%d N=%d nvec=%d:", nR, ntot, nvec);
  std::string rec = line;
  for (int r : rows_opt)
    for (int k : ks_opt) {
      double* part = c->w_part.get<double>((size_t)k * nR * nvec);
      GemvArgs a{};
      a.C = C;
      a.ld = ld;
      a.ntot = ntot;
      a.F0 = F0;
      a.nF = nR;
      a.G = G;
      a.gstride = gstride;
      a.nvec = nvec;
      a.partial = part;
      a.ksplit = k;
      a.rows = r;
      launch_gemv(a, c->stream);
      GAIN_HIP(hipEventRecord(e0, c->stream));
      for (int q = 0; q < 3; ++q) launch_gemv(a, c->stream);
      GAIN_HIP(hipEventRecord(e1, c->stream));
      GAIN_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      GAIN_HIP(hipEventElapsedTime(&ms, e0, e1));
      ms /= 3.f;
      snprintf(line, sizeof(line), " %dx%d=%.3fms", r, k, ms);
      rec += line;
      if (ms < best) {
        best = ms;
        br = r;
        bk = k;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  snprintf(line, sizeof(line), " -> rows %d ksplit %d\n", br, bk);
  c->tune_log += rec + line;
  c->gemv_rows = br;
  c->gemv_ksplit = bk;
  c->tuned[key] = br * 100 + bk;
}

// Lines for nI initial states of one block against one final block.
void run_lines(gain_ctx* c, const Plan& P, const double* cI_host,
               int precision, bool plain, const gain_lines* out) {
  Basis& bF = *P.bF;
  Block& blkF = *P.blkF;
  const int nI = P.nI, nF = blkF.n;
  const int md2 = c->maxdeg * c->maxdeg;
  const Staged st = stage(c, P, cI_host);
  const Small& sm = st.sm;
  const double* cbase = st.cbase;
  const int nvpi = st.nvpi, nvecX = st.nvecX, nvecG = st.nvecG;
  const bool compute = st.compute;

  const int nR = P.Fhi - P.Flo;
  const Irrep& irF = bF.irrep[P.gF];
  const int ntotF = irF.ntot;
  const int64_t gstride = (ntotF + 1) / 2 * 2;
  const double* src = nullptr;
  int ksplit = 1;
  int64_t src_ks = 0, src_f = 0, src_v = 0;
  if (compute) {
    const int64_t hstride = bF.nprim;
    double* H = c->w_H.get<double>((size_t)(nvecX + 1) * hstride);
    launch_fill_zero(H + (size_t)nvecX * hstride, hstride, c->stream);
    run_half_ls(c, P, cbase, sm, nvecX, H, plain ? GAIN_PREC_FP64 : precision);
    double* G = c->w_G.get<double>((size_t)nvecG * gstride);
    {
      Timed t(c, "expand_T");
      launch_expand_T(make_expand_T_args(P, H, sm, nvecG, G, gstride),
                      c->stream);
    }
    const double* Crows = blkF.vec;
    if (plain && nI == 1) {
      if (c->autotune) gemv_tune(c, Crows, blkF.ld, ntotF, P.Flo, nR, G,
                                 gstride, nvpi);
      ksplit = c->gemv_ksplit;
      double* part = c->w_part.get<double>((size_t)ksplit * nR * nvpi);
      Timed t(c, "gemv");
      GemvArgs a{};
      a.C = Crows;
      a.ld = blkF.ld;
      a.ntot = ntotF;
      a.F0 = P.Flo;
      a.nF = nR;
      a.G = G;
      a.gstride = gstride;
      a.nvec = nvpi;
      a.partial = part;
      a.ksplit = ksplit;
      a.rows = c->gemv_rows;
      launch_gemv(a, c->stream);
      c->last["gemv_grid"] = gemv_grid_items(a);
      src = part;
      src_ks = (int64_t)nR * nvpi;
      src_f = nvpi;
      src_v = 1;
    } else {
      double* LS = c->w_part.get<double>((size_t)nR * nvecG);
      const int S = slices_of(precision);
      const int S8 = i8_slices_of(precision);
      // split-INT8: final records sliced once per block, projected vectors
      // sliced here, outside the timed projection GEMM
      // CRT form: the projection may use its own moduli count (tuning key
      // "proj_crt_moduli"; the half line strength keeps the precision's)
      const int NMc = crt_moduli_of(precision) > 0 && c->proj_crt > 0
                          ? c->proj_crt
                          : crt_moduli_of(precision);
      const Sliced* cs8 = nullptr;
      int8_t* gs8 = nullptr;
      int* ge8 = nullptr;
      int64_t gplane8 = 0;
      if (S8 > 0 || NMc > 0) {
        const int np = S8 > 0 ? S8 : NMc;
        cs8 = S8 > 0 ? &block_sliced_i8(c, blkF, S8)
                     : &block_sliced_crt(c, blkF, NMc);
        const int64_t Rg = i8_rows_pad(nvecG);
        gplane8 = Rg * cs8->Kp;
        gs8 = c->w_gs.get<int8_t>((size_t)np * gplane8);
        ge8 = c->w_ge.get<int>((size_t)Rg);
        if (S8 > 0)
          slice_rows_i8(S8, G, gstride, nvecG, ntotF, gs8, gplane8, ge8,
                        c->stream);
        else
          slice_rows_crt(NMc, cs8->beta, G, gstride, nvecG, ntotF, gs8,
                         gplane8, ge8, c->stream);
      }
      Timed t(c, "gemm_proj");
      if (S > 0) {
        const Sliced& cs = block_sliced(c, blkF, S);
        uint16_t* gs = c->w_gs.get<uint16_t>((size_t)S * nvecG * cs.Kp);
        const int nch = (int)slice_nchunk(S, ntotF);
        int* ge = c->w_ge.get<int>((size_t)nvecG * nch);
        slice_rows_bf16(S, G, gstride, nvecG, ntotF, gs, ge, c->stream);
        SlicedGemmArgs g{};
        g.S = S;
        g.M = nR;
        g.N = nvecG;
        g.Kp = cs.Kp;
        g.A = cs.planes + (int64_t)P.Flo * cs.Kp;
        g.a_plane = (int64_t)blkF.n * cs.Kp;
        g.eA = cs.expo + (int64_t)P.Flo * nch;
        g.nch = nch;
        g.B = gs;
        g.b_plane = (int64_t)nvecG * cs.Kp;
        g.eB = ge;
        g.C = LS;
        g.ldc = nR;
        g.nz = 1;
        {
          Timed tg(c, "proj_gemm");
          gemm_sliced_bf16(g, c->stream);
        }
      } else if (NMc > 0) {
        const Sliced& cs = *cs8;
        I8GemmArgs g{};
        g.S = NMc;
        g.beta = cs.beta;
        g.M = nR;
        g.N = nvecG;
        g.row0 = P.Flo;
        g.Kp = cs.Kp;
        g.A = cs.p8;
        g.a_plane = cs.plane;
        g.eA = cs.expo;
        g.B = gs8;
        g.b_plane = gplane8;
        g.eB = ge8;
        g.C = LS;
        g.ldc = nR;
        g.nz = 1;
        g.R = crt_residues(c, g);
        g.cfg = crt_cfg_of(c, 0, g);
        {
          Timed tg(c, "proj_gemm");
          gemm_crt_res(g, c->stream);
        }
        gemm_crt_combine(g, c->stream);
        c->last["proj_cfg"] = g.cfg;
        c->last["proj_split"] = 1;
        c->last["proj_grid"] = gemm_crt_grid_items(g);
      } else if (S8 > 0) {
        const Sliced& cs = *cs8;
        I8GemmArgs g{};
        g.S = S8;
        g.M = nR;
        g.N = nvecG;
        g.row0 = P.Flo;
        g.Kp = cs.Kp;
        g.A = cs.p8;
        g.a_plane = cs.plane;
        g.a_z = 0;
        g.eA = cs.expo;
        g.ea_z = 0;
        g.B = gs8;
        g.b_plane = gplane8;
        g.eB = ge8;
        g.ldc = nR;
        g.c_z = 0;
        g.c_ks = (int64_t)nR * nvecG;
        g.nz = 1;
        // the tuner may write up to 4 split partials
        LS = c->w_part.get<double>((size_t)4 * nR * nvecG);
        g.C = LS;
        int ks = 1;
        g.cfg = i8_cfg_split(c, 0, g, 4, &ks);
        g.ksplit = ks;
        {
          Timed tg(c, "proj_gemm");
          gemm_sliced_i8(g, c->stream);
        }
        c->last["proj_cfg"] = g.cfg;
        c->last["proj_split"] = ks;
        c->last["proj_grid"] = gemm_i8_grid_items(g);
        ksplit = ks;
        src_ks = (int64_t)nR * nvecG;
      } else if (c->use_mfma && precision == GAIN_PREC_FP64) {
        GemmArgs g{};
        g.M = nR;
        g.N = nvecG;
        g.K = ntotF;
        g.A = Crows + (int64_t)P.Flo * blkF.ld;
        g.lda = blkF.ld;
        g.a_kcontig = true;
        g.B = G;
        g.ldb = gstride;
        g.ldc = nR;
        g.sCks = (int64_t)nR * nvecG;
        g.nz = 1;
        g.ksplit = 1;
        // the tuner may write every split count: room for the largest
        LS = c->w_part.get<double>((size_t)8 * nR * nvecG);
        g.C = LS;
        int ks = 1;
        g.cfg = gemm_cfg_split(c, 0, g, &ks, true);
        g.ksplit = ks;
        {
          Timed tg(c, "proj_gemm");
          gemm_f64(g, c->stream);
        }
        c->last["proj_cfg"] = g.cfg < 0 ? gemm_default_config(g.N) : g.cfg;
        c->last["proj_split"] = ks;
        c->last["proj_grid"] = gemm_grid_items(g);
        ksplit = ks;
        src_ks = (int64_t)nR * nvecG;
      } else if (precision != GAIN_PREC_FP64) {
        Timed tg(c, "proj_gemm");
        gemm_projection(precision, nR, nvecG, ntotF,
                        Crows + (int64_t)P.Flo * blkF.ld, blkF.ld, G, gstride,
                        LS, nR, c->stream, c->f32_cfg);
      } else
        launch_gemm_tn_valu(nR, nvecG, ntotF,
                            Crows + (int64_t)P.Flo * blkF.ld, blkF.ld, G,
                            gstride, LS, nR, c->stream);
      src = LS;
      src_f = 1;
      src_v = nR;
    }
  }
  // ---- epilogue
  const bool host_out = !out->on_device;
  const size_t nPairs = (size_t)nI * nF;
  double *o_ls = out->ls, *o_S = out->S, *o_A = out->A, *o_I = out->I;
  uint8_t* o_fl = out->flags;
  char* stage = nullptr;
  size_t s_ls = 0, s_S = 0, s_A = 0, s_I = 0, s_fl = 0, s_tot = 0;
  if (host_out) {
    s_ls = 0;
    s_S = s_ls + (out->ls ? nPairs * md2 * sizeof(double) : 0);
    s_A = s_S + (out->S ? nPairs * sizeof(double) : 0);
    s_I = s_A + (out->A ? nPairs * sizeof(double) : 0);
    s_fl = s_I + (out->I ? nPairs * sizeof(double) : 0);
    s_tot = s_fl + (out->flags ? nPairs : 0);
    stage = c->w_out.get<char>(s_tot);
    o_ls = out->ls ? (double*)(stage + s_ls) : nullptr;
    o_S = out->S ? (double*)(stage + s_S) : nullptr;
    o_A = out->A ? (double*)(stage + s_A) : nullptr;
    o_I = out->I ? (double*)(stage + s_I) : nullptr;
    o_fl = out->flags ? (uint8_t*)(stage + s_fl) : nullptr;
  }
  {
    Timed t(c, "epilogue");
    LinesArgs la = make_lines_args(c, P, sm, nvpi, compute);
    la.src = src;
    la.ksplit = ksplit;
    la.src_ks = src_ks;
    la.src_f = src_f;
    la.src_v = src_v;
    la.ls = o_ls;
    la.S = o_S;
    la.A = o_A;
    la.I = o_I;
    la.flags = o_fl;
    launch_lines(la, c->stream);
  }
  if (host_out) {
    if (out->ls)
      GAIN_HIP(hipMemcpyAsync(out->ls, o_ls, nPairs * md2 * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
    if (out->S)
      GAIN_HIP(hipMemcpyAsync(out->S, o_S, nPairs * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
    if (out->A)
      GAIN_HIP(hipMemcpyAsync(out->A, o_A, nPairs * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
    if (out->I)
      GAIN_HIP(hipMemcpyAsync(out->I, o_I, nPairs * sizeof(double),
                              hipMemcpyDeviceToHost, c->stream));
    if (out->flags)
      GAIN_HIP(hipMemcpyAsync(out->flags, o_fl, nPairs, hipMemcpyDeviceToHost,
                              c->stream));
    GAIN_HIP(hipStreamSynchronize(c->stream));
  }
}


// Adjoint of the batched FP64 path: recomputes the forward, then
//   dls = dS * dS/dls,  dG = C_F^T dls,  dH = P_F dG,  dY = gather^T dH,
//   dmu_alpha = dY_alpha X^T  (a rank-k update, k = columns of X).
void run_adjoint(gain_ctx* c, const Plan& P, const double* dS_dev,
                 double* dmu_work, const double* ls_fwd,
                 const uint8_t* flags_fwd) {
  const gain_params& p = c->params;
  Basis& bI = *P.bI;
  Basis& bF = *P.bF;
  Block& blkF = *P.blkF;
  const int nI = P.nI, nF = blkF.n, nc = c->nc;
  if (bI.rotsym)
    fail(GAIN_ERR_ARG, "the adjoint covers K-block bases only (J index %d "
         "is a symmetric-top basis)", P.jindI);
  const Staged st = stage(c, P, nullptr);
  const Small& sm = st.sm;
  launch_fill_zero(dmu_work, (int64_t)3 * nc * nc, c->stream);
  if (!st.compute) return;
  const int nvpi = st.nvpi, nvecX = st.nvecX, nvecG = st.nvecG;
  const int nR = P.Fhi - P.Flo;
  const Irrep& irF = bF.irrep[P.gF];
  const int ntotF = irF.ntot;
  const int64_t gstride = (ntotF + 1) / 2 * 2;
  const int64_t hstride = bF.nprim;
  const double* Crows = blkF.vec + (int64_t)P.Flo * blkF.ld;
  LinesArgs la = make_lines_args(c, P, sm, nvpi, true);
  // adjoint work buffers: LS and dLS share w_part, flags in w_out
  double* LS = c->w_part.get<double>((size_t)2 * nR * nvecG);
  double* dLS = LS + (size_t)nR * nvecG;
  const double* X = nullptr;
  const uint8_t* flags = flags_fwd;
  if (bF.nvib > nc)
    fail(GAIN_ERR_ARG,
         "vibrational index %d of J=%d exceeds the dipole dimension %d",
         bF.nvib - 1, bF.J, nc);
  if (ls_fwd) {
    // the forward's line strengths and flags are given: only X (the
    // scattered initial vectors of the rank-k update) is rebuilt
    ensure_slots(c, bI);
    X = c->w_X.get<double>((size_t)nvecX * bI.ncol * nc);
    launch_scatter_expand(make_scatter_args(c, P, st.cbase, sm, nvecX,
                                            (double*)X),
                          c->stream);
  } else {
    // ---- forward (FP64)
    double* H = c->w_H.get<double>((size_t)(nvecX + 1) * hstride);
    launch_fill_zero(H + (size_t)nvecX * hstride, hstride, c->stream);
    run_half_ls(c, P, st.cbase, sm, nvecX, H, GAIN_PREC_FP64);
    X = (const double*)c->w_X.p;
    double* G = c->w_G.get<double>((size_t)nvecG * gstride);
    launch_expand_T(make_expand_T_args(P, H, sm, nvecG, G, gstride),
                    c->stream);
    gemm_projection(GAIN_PREC_FP64, nR, nvecG, ntotF, Crows, blkF.ld, G,
                    gstride, LS, nR, c->stream);
    uint8_t* fl = c->w_out.get<uint8_t>((size_t)nI * nF);
    la.src = LS;
    la.src_f = 1;
    la.src_v = nR;
    la.flags = fl;
    launch_lines(la, c->stream);
    flags = fl;
  }
  // ---- backward
  Timed t(c, "adjoint");
  const int ndegI = p.sym_degen[P.gI], ndegF = p.sym_degen[P.gF];
  DlsArgs da{};
  da.nI = nI;
  da.nF = nF;
  da.Flo = P.Flo;
  da.nR = nR;
  da.nvpi = nvpi;
  da.dS = dS_dev;
  da.flags = flags;
  da.eI = sm.eI;
  da.eF = blkF.d_energy;
  da.ls = LS;
  da.scale = (p.reduced && ndegF != 1 && ndegI != 1) ? 2.0
                                                     : 2.0 / (double)ndegI;
  da.dls = dLS;
  if (ls_fwd) {
    da.ls_pair = ls_fwd;
    da.slot_of_vec = la.slot_of_vec;
    da.md2 = c->maxdeg * c->maxdeg;
  }
  launch_dls(da, c->stream);
  // dG[v][t] = sum_f C[Flo+f][t] dLS[v][f]
  double* dG = c->w_G.get<double>((size_t)nvecG * gstride);
  {
    GemmArgs g{};
    g.M = ntotF;
    g.N = nvecG;
    g.K = nR;
    g.A = Crows;
    g.lda = blkF.ld;
    g.a_kcontig = false;
    g.B = dLS;
    g.ldb = nR;
    g.C = dG;
    g.ldc = gstride;
    g.nz = 1;
    g.ksplit = 1;
    gemm_f64(g, c->stream);
  }
  double* dH = c->w_H.get<double>((size_t)(nvecX + 1) * hstride);
  {
    ExpandAccArgs xa{};
    xa.nprim = bF.nprim;
    xa.ntot = ntotF;
    xa.sdeg = irF.sdeg;
    xa.icontr = bF.icontr;
    xa.ijt = bF.ijterms + (int64_t)P.gF * bF.maxsymcoeffs;
    xa.N = irF.N;
    xa.repres = irF.repres;
    xa.dG = dG;
    xa.gstride = gstride;
    xa.vptr = sm.vptr;
    xa.vidx = sm.vidx;
    xa.vdeg = sm.vdeg;
    xa.nH = nvecX;
    xa.dH = dH;
    xa.hstride = hstride;
    launch_expand_acc(xa, c->stream);
  }
  ensure_slots(c, bF);
  const int ncolTot = nvecX * bI.ncol;
  double* dY = c->w_Y.get<double>((size_t)3 * ncolTot * nc);
  {
    GatherTArgs ga{};
    ga.nc = nc;
    ga.jI = bI.J;
    ga.jF = bF.J;
    ga.jmax = c->jmax;
    ga.ncolI = bI.ncol;
    ga.ncolF = bF.ncol;
    ga.nvec = nvecX;
    ga.slot_ptrF = bF.slot_ptr;
    ga.slot_primF = bF.slot_prim;
    ga.threej = c->d_threej;
    ga.dH = dH;
    ga.hstride = hstride;
    ga.dY = dY;
    launch_gather_T(ga, c->stream);
  }
  double* XT = c->w_tmp.get<double>((size_t)ncolTot * nc);
  launch_transpose(X, ncolTot, nc, XT, c->stream);
  Timed tr(c, "adjoint_rankk");
  GemmArgs g{};
  g.M = nc;
  g.N = nc;
  g.K = ncolTot;
  g.A = dY;
  g.lda = nc;
  g.sAz = (int64_t)ncolTot * nc;
  g.a_kcontig = false;
  g.B = XT;
  g.ldb = ncolTot;
  g.C = dmu_work;
  g.ldc = nc;
  g.sCz = (int64_t)nc * nc;
  g.nz = 3;
  g.ksplit = 1;
  gemm_f64(g, c->stream);
}
}  // namespace gain

using namespace gain;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" const char* gain_last_error(void) { return t_last_error.c_str(); }

extern "C" const char* gain_version(void) { return "gain_engine 0.1 gfx950"; }

extern "C" int gain_create(gain_ctx** out, int device, int jmax, int nrepres,
                           int maxdegen) {
  return guarded([&] {
    GAIN_CHECK_ARG(out, "null output pointer");
    *out = nullptr;
    GAIN_CHECK_ARG(jmax >= 0 && jmax <= 200, "jmax %d out of range", jmax);
    GAIN_CHECK_ARG(nrepres >= 1 && nrepres <= GAIN_MAX_IRREPS,
                   "nrepres %d out of range", nrepres);
    GAIN_CHECK_ARG(maxdegen >= 1 && maxdegen <= 3, "maxdegen %d out of range",
                   maxdegen);
    int ndev = 0;
    GAIN_HIP(hipGetDeviceCount(&ndev));
    GAIN_CHECK_ARG(device >= 0 && device < ndev, "device %d not present (%d)",
                   device, ndev);
    gain_ctx* c = new gain_ctx();
    try {
      c->device = device;
      DeviceGuard g(device);
      GAIN_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
      c->jmax = jmax;
      c->nrepres = nrepres;
      c->maxdeg = maxdegen;
      const int n1 = jmax + 1;
      c->threej.assign((size_t)n1 * n1 * 9, 0.0);
      for (int jI = 0; jI <= jmax; ++jI)
        for (int jF = std::max(jI - 1, 0); jF <= std::min(jI + 1, jmax); ++jF)
          for (int kI = 0; kI <= jI; ++kI)
            for (int kF = std::max(kI - 1, 0); kF <= std::min(kI + 1, jF);
                 ++kF)
              c->threej[jI + kI * n1 + (jF - jI + 1) * n1 * n1 +
                        (kF - kI + 1) * n1 * n1 * 3] =
                  three_j(jI, 1, jF, kI, kF - kI, -kF);
      c->d_threej = c->arena.upload(c->threej.data(), c->threej.size());
      noalloc_init_device(device);
    } catch (...) {
      if (c->stream) (void)hipStreamDestroy(c->stream);
      delete c;
      throw;
    }
    *out = c;
  });
}

extern "C" int gain_destroy(gain_ctx* c) {
  return guarded([&] {
    if (!c) return;
    {
      DeviceGuard g(c->device);
      (void)hipStreamSynchronize(c->stream);
      for (auto& kv : c->timers) {
        for (auto& e : kv.second.pending) {
          (void)hipEventDestroy(e.first);
          (void)hipEventDestroy(e.second);
        }
        for (auto& e : kv.second.pool) {
          (void)hipEventDestroy(e.first);
          (void)hipEventDestroy(e.second);
        }
      }
      (void)hipStreamDestroy(c->stream);
    }
    delete c;
  });
}

extern "C" int gain_get_threej(gain_ctx* c, double* host_out) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(host_out, "null output");
    std::memcpy(host_out, c->threej.data(), sizeof(double) * c->threej.size());
  });
}

extern "C" int gain_get_stream(gain_ctx* c, void** s) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(s, "null output");
    *s = (void*)c->stream;
  });
}

extern "C" int gain_synchronize(gain_ctx* c) {
  return guarded([&] {
    need(c);
    DeviceGuard g(c->device);
    GAIN_HIP(hipStreamSynchronize(c->stream));
  });
}

extern "C" int gain_set_params(gain_ctx* c, const gain_params* p) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(p, "null params");
    GAIN_CHECK_ARG(p->nrepres == c->nrepres, "nrepres %d != context's %d",
                   p->nrepres, c->nrepres);
    GAIN_CHECK_ARG(p->maxdeg == c->maxdeg, "maxdeg %d != context's %d",
                   p->maxdeg, c->maxdeg);
    GAIN_CHECK_ARG(p->temperature > 0.0 && p->partition != 0.0,
                   "temperature and partition function must be set");
    for (int g = 0; g < p->nrepres; ++g)
      GAIN_CHECK_ARG(p->sym_degen[g] >= 1 && p->sym_degen[g] <= p->maxdeg,
                     "degeneracy of irrep %d out of range", g);
    // TroveInput::ComputeIGammaPair, TroveInput.cpp:202-236
    int pair[GAIN_MAX_IRREPS];
    for (int gI = 0; gI < p->nrepres; ++gI) {
      int ng = 0;
      pair[gI] = gI;
      for (int gF = 0; gF < p->nrepres; ++gF)
        if (gI != gF && p->isym_pairs[gI] == p->isym_pairs[gF]) {
          pair[gI] = gF;
          if (++ng > 1)
            fail(GAIN_ERR_CONFIG,
                 "selection rules do not come in pairs (irrep %d)", gI);
        }
      if (p->gns[gI] != p->gns[pair[gI]])
        fail(GAIN_ERR_CONFIG, "selection rules do not agree with gns "
                              "(irrep %d)", gI);
    }
    c->params = *p;
    std::memcpy(c->pair, pair, sizeof(pair));
    c->params_set = true;
  });
}

extern "C" int gain_get_gamma_pair(gain_ctx* c, int gamma, int* pair) {
  return guarded([&] {
    need(c);
    if (!c->params_set) fail(GAIN_ERR_STATE, "gain_set_params not called");
    GAIN_CHECK_ARG(gamma >= 0 && gamma < c->nrepres && pair, "bad irrep");
    *pair = c->pair[gamma];
  });
}

extern "C" int gain_upload_dipole(gain_ctx* c, const double* mu, int nc,
                                  int startF, int endF) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(mu && nc > 0, "bad dipole");
    GAIN_CHECK_ARG(startF >= 0 && startF < endF && endF <= nc,
                   "dipole piece [%d, %d) outside [0, %d)", startF, endF, nc);
    DeviceGuard g(c->device);
    ++c->gen;
    release_sliced(c, c->mu_sl);
    if (startF == 0 && endF == nc) {  // the whole matrix
      if (c->d_mu) c->arena.release(c->d_mu);
      c->d_mu = nullptr;
      c->d_mu = c->arena.upload(mu, (size_t)3 * nc * nc);
      c->nc = nc;
      c->mu_have.assign(nc, 1);
      c->mu_missing = 0;
      build_slots_all(c);
      return;
    }
    // a piece of final columns [startF, endF) (the reference transfers the
    // dipole in such pieces when it does not fit, TroveDipole.cpp:47-112):
    // copied into the resident matrix [alpha][vI][vF]
    if (!c->d_mu || c->nc != nc) {
      if (c->d_mu) c->arena.release(c->d_mu);
      c->d_mu = nullptr;
      c->d_mu = c->arena.alloc<double>((size_t)3 * nc * nc);
      c->nc = nc;
      c->mu_have.assign(nc, 0);
      c->mu_missing = nc;
      build_slots_all(c);
    }
    const int w = endF - startF;
    for (int a = 0; a < 3; ++a)
      GAIN_HIP(hipMemcpy2D(c->d_mu + (size_t)a * nc * nc + startF,
                           sizeof(double) * nc, mu + (size_t)a * w * nc,
                           sizeof(double) * w, sizeof(double) * w, nc,
                           hipMemcpyHostToDevice));
    for (int v = startF; v < endF; ++v)
      if (!c->mu_have[v]) {
        c->mu_have[v] = 1;
        --c->mu_missing;
      }
  });
}

extern "C" int gain_upload_basis(gain_ctx* c, int jind, int J, int nprim,
                                 const int* kstart, const int* kblock,
                                 const int* ktau, const int* vib) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(jind >= 0 && jind < 1024, "J index out of range");
    GAIN_CHECK_ARG(J >= 0 && J <= c->jmax, "J=%d above jmax %d", J, c->jmax);
    GAIN_CHECK_ARG(nprim > 0 && kstart && kblock && ktau && vib,
                   "bad basis arrays");
    // K blocks must tile [0, nprim) in order (TroveBasisSet.cpp:658-681)
    int pos = 0;
    for (int K = 0; K <= J; ++K) {
      GAIN_CHECK_ARG(kstart[K] == pos && kblock[K] >= 0,
                     "K block %d does not start at %d", K, pos);
      pos += kblock[K];
    }
    GAIN_CHECK_ARG(pos == nprim, "K blocks cover %d of %d primitives", pos,
                   nprim);
    DeviceGuard g(c->device);
    if ((int)c->bases.size() <= jind) c->bases.resize(jind + 1);
    Basis& b = c->bases[jind];
    if (b.present) fail(GAIN_ERR_STATE, "J index %d already uploaded", jind);
    b.J = J;
    b.nprim = nprim;
    b.ncol = 2 * (J + 1);
    b.kstart.assign(kstart, kstart + J + 1);
    b.kblock.assign(kblock, kblock + J + 1);
    b.hk.resize(nprim);
    b.htau.assign(ktau, ktau + nprim);
    b.hvib.assign(vib, vib + nprim);
    int nv = 0;
    for (int K = 0; K <= J; ++K)
      for (int r = kstart[K]; r < kstart[K] + kblock[K]; ++r) b.hk[r] = K;
    b.hcol.resize(nprim);
    for (int r = 0; r < nprim; ++r) {
      GAIN_CHECK_ARG(vib[r] >= 0, "negative vibrational index");
      nv = std::max(nv, vib[r] + 1);
      b.htau[r] &= 1;
      b.hcol[r] = 2 * b.hk[r] + b.htau[r];
    }
    b.nvib = nv;
    b.kval = c->arena.upload(b.hk.data(), nprim);
    b.tau = c->arena.upload(b.htau.data(), nprim);
    b.vib = c->arena.upload(b.hvib.data(), nprim);
    b.present = true;
    if (c->d_mu && b.nvib <= c->nc) ensure_slots(c, b);
  });
}

extern "C" int gain_upload_basis_rotsym(gain_ctx* c, int jind, int J,
                                        int nprim, int nlevels, int maxdeg,
                                        const int* level, const int* deg,
                                        const int* vib) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(jind >= 0 && jind < 1024, "J index out of range");
    GAIN_CHECK_ARG(J >= 0 && J <= c->jmax, "J=%d above jmax %d", J, c->jmax);
    GAIN_CHECK_ARG(nprim > 0 && level && deg && vib, "bad basis arrays");
    GAIN_CHECK_ARG(nlevels >= 1 && maxdeg >= 1 && maxdeg <= 3,
                   "bad level count %d / degeneracy %d", nlevels, maxdeg);
    DeviceGuard g(c->device);
    if ((int)c->bases.size() <= jind) c->bases.resize(jind + 1);
    Basis& b = c->bases[jind];
    if (b.present) fail(GAIN_ERR_STATE, "J index %d already uploaded", jind);
    b.J = J;
    b.nprim = nprim;
    b.rotsym = true;
    b.nlevels = nlevels;
    b.rmd = maxdeg;
    b.ncol = nlevels * maxdeg;
    b.hk.assign(nprim, 0);  // no K-block sign on the initial side
    b.htau.assign(nprim, 0);
    b.hvib.assign(vib, vib + nprim);
    b.hcol.resize(nprim);
    int nv = 0;
    for (int r = 0; r < nprim; ++r) {
      GAIN_CHECK_ARG(level[r] >= 0 && level[r] < nlevels && deg[r] >= 0 &&
                         deg[r] < maxdeg && vib[r] >= 0,
                     "primitive %d: level %d / component %d / vib %d out of "
                     "range", r, level[r], deg[r], vib[r]);
      b.hcol[r] = level[r] * maxdeg + deg[r];
      nv = std::max(nv, vib[r] + 1);
    }
    b.nvib = nv;
    b.kval = c->arena.upload(b.hk.data(), nprim);
    b.tau = c->arena.upload(b.htau.data(), nprim);
    b.vib = c->arena.upload(b.hvib.data(), nprim);
    b.present = true;
    if (c->d_mu && b.nvib <= c->nc) ensure_slots(c, b);
  });
}

extern "C" int gain_upload_wigner(gain_ctx* c, int jindI, int jindF,
                                  int nlevelsI, int nlevelsF, int maxdeg,
                                  const double* rot) {
  return guarded([&] {
    need(c);
    Basis& bI = basis_of(c, jindI, false);
    Basis& bF = basis_of(c, jindF, false);
    GAIN_CHECK_ARG(bI.rotsym && bF.rotsym,
                   "rotational tables belong to symmetric-top bases");
    GAIN_CHECK_ARG(std::abs(bF.J - bI.J) <= 1, "|J_F - J_I| > 1");
    GAIN_CHECK_ARG(rot && nlevelsI == bI.nlevels && nlevelsF == bF.nlevels &&
                       maxdeg == bI.rmd && maxdeg == bF.rmd,
                   "table %d x %d levels, degeneracy %d; bases have %d x %d, "
                   "%d / %d", nlevelsI, nlevelsF, maxdeg, bI.nlevels,
                   bF.nlevels, bI.rmd, bF.rmd);
    DeviceGuard g(c->device);
    auto it = c->wigner.find({jindI, jindF});
    if (it != c->wigner.end()) {
      ++c->gen;  // a sweep may hold the old table
      c->arena.release(it->second.d);
      c->wigner.erase(it);
    }
    WignerTable w;
    w.nlI = nlevelsI;
    w.nlF = nlevelsF;
    w.md = maxdeg;
    w.d = c->arena.upload(rot, (size_t)3 * nlevelsI * nlevelsF * maxdeg *
                                   maxdeg);
    c->wigner[{jindI, jindF}] = w;
  });
}

extern "C" int gain_upload_inflation(gain_ctx* c, int jind, int maxsymcoeffs,
                                     const int* ijterms, const int* icontr,
                                     int matsize, const int* ntot,
                                     const int* sdeg, const int** N,
                                     const double** repres) {
  return guarded([&] {
    need(c);
    Basis& b = basis_of(c, jind, false);
    GAIN_CHECK_ARG(maxsymcoeffs > 0 && ijterms && icontr && matsize > 0 &&
                       ntot && sdeg && N && repres,
                   "bad inflation arrays");
    if (b.has_infl) fail(GAIN_ERR_STATE, "inflation of J index %d already set",
                         jind);
    DeviceGuard g(c->device);
    for (int r = 0; r < b.nprim; ++r) {
      GAIN_CHECK_ARG(icontr[r] >= 0 && icontr[r] < maxsymcoeffs,
                     "icontr case out of range at %d", r);
      GAIN_CHECK_ARG(icontr[r + b.nprim] >= 0 && icontr[r + b.nprim] < matsize,
                     "icontr lambda out of range at %d", r);
    }
    b.maxsymcoeffs = maxsymcoeffs;
    b.matsize = matsize;
    b.ijterms = c->arena.upload(ijterms, (size_t)maxsymcoeffs * c->nrepres);
    b.icontr = c->arena.upload(icontr, (size_t)2 * b.nprim);
    // case -> primitives CSR (for the transposed expansion)
    std::vector<int> cp(maxsymcoeffs + 1, 0), prim(b.nprim);
    for (int r = 0; r < b.nprim; ++r) cp[icontr[r] + 1]++;
    for (int k = 0; k < maxsymcoeffs; ++k) cp[k + 1] += cp[k];
    std::vector<int> fill(cp.begin(), cp.end() - 1);
    for (int r = 0; r < b.nprim; ++r) prim[fill[icontr[r]]++] = r;
    b.case_ptr = c->arena.upload(cp.data(), cp.size());
    b.case_prim = c->arena.upload(prim.data(), prim.size());
    for (int gm = 0; gm < c->nrepres; ++gm) {
      Irrep& ir = b.irrep[gm];
      ir.ntot = ntot[gm];
      ir.sdeg = sdeg[gm];
      GAIN_CHECK_ARG(ir.ntot >= 0 && ir.sdeg >= 1 && ir.sdeg <= c->maxdeg,
                     "irrep %d: bad Ntot/sdeg", gm);
      const int* ijt = ijterms + (size_t)gm * maxsymcoeffs;
      std::vector<int> tcase(std::max(ir.ntot, 1), 0);
      for (int k = 0; k < maxsymcoeffs; ++k) {
        GAIN_CHECK_ARG(N[gm][k] >= 0 && ijt[k] >= 0 &&
                           ijt[k] + N[gm][k] <= ir.ntot,
                       "irrep %d case %d: ijterms/N exceed Ntot", gm, k);
        for (int i = 0; i < N[gm][k]; ++i) tcase[ijt[k] + i] = k;
      }
      ir.N = c->arena.upload(N[gm], maxsymcoeffs);
      ir.maxN = 0;
      for (int k = 0; k < maxsymcoeffs; ++k)
        ir.maxN = std::max(ir.maxN, N[gm][k]);
      const size_t nrep = (size_t)ir.ntot * ir.sdeg * matsize;
      ir.repres = nrep ? c->arena.upload(repres[gm], nrep)
                       : c->arena.alloc<double>(1);
      ir.tcase = c->arena.upload(tcase.data(), tcase.size());
    }
    b.has_infl = true;
  });
}

static void register_block(gain_ctx* c, int jind, int gamma, int nstates,
                           int rec_len, int64_t ld, const double* energies,
                           const double* vec, bool device) {
  need(c);
  GAIN_CHECK_ARG(gamma >= 0 && gamma < c->nrepres, "irrep out of range");
  GAIN_CHECK_ARG(nstates >= 0 && rec_len >= 0 && ld >= rec_len,
                 "bad block sizes");
  GAIN_CHECK_ARG(energies || nstates == 0, "null energies");
  DeviceGuard g(c->device);
  auto it = c->blocks.find({jind, gamma});
  if (it != c->blocks.end()) {
    ++c->gen;
    release_sliced(c, it->second.sl);
    if (it->second.owned && it->second.vec)
      c->arena.release((void*)it->second.vec);
    if (it->second.d_energy) c->arena.release(it->second.d_energy);
    c->blocks.erase(it);
  }
  Block b;
  b.n = nstates;
  b.rec_len = rec_len;
  b.ld = ld;
  b.energy.assign(energies, energies + nstates);
  for (int i = 1; i < nstates; ++i)
    if (!(b.energy[i - 1] <= b.energy[i])) b.sorted = false;
  b.d_energy = c->arena.upload(b.energy.data(), nstates);
  if (vec) {
    if (device) {
      b.vec = vec;
      b.owned = false;
    } else {
      double* d = c->arena.alloc<double>((size_t)nstates * ld);
      GAIN_HIP(hipMemcpy(d, vec, sizeof(double) * nstates * ld,
                         hipMemcpyHostToDevice));
      b.vec = d;
      b.owned = true;
    }
  }
  c->blocks[{jind, gamma}] = std::move(b);
}

extern "C" int gain_upload_block(gain_ctx* c, int jind, int gamma,
                                 int nstates, int rec_len,
                                 const double* energies,
                                 const double* vectors) {
  return guarded([&] {
    register_block(c, jind, gamma, nstates, rec_len, rec_len, energies,
                   vectors, false);
  });
}

extern "C" int gain_bind_block_device(gain_ctx* c, int jind, int gamma,
                                      int nstates, int rec_len, int64_t ld,
                                      const double* energies,
                                      const double* vectors_device) {
  return guarded([&] {
    GAIN_CHECK_ARG(vectors_device, "null device pointer");
    register_block(c, jind, gamma, nstates, rec_len, ld, energies,
                   vectors_device, true);
  });
}

extern "C" int gain_half_ls(gain_ctx* c, int jindI, int gammaI, int idegI,
                            const double* cI, int jindF, double* h_out) {
  return guarded([&] {
    need(c);
    need_mu(c);
    GAIN_CHECK_ARG(cI && h_out, "null vector");
    GAIN_CHECK_ARG(gammaI >= 0 && gammaI < c->nrepres, "irrep out of range");
    DeviceGuard g(c->device);
    Plan P;
    P.jindI = jindI;
    P.gI = gammaI;
    P.jindF = jindF;
    P.gF = gammaI;
    P.bI = &basis_of(c, jindI, true);
    P.bF = &basis_of(c, jindF, false);
    GAIN_CHECK_ARG(std::abs(P.bI->J - P.bF->J) <= 1, "|J_F - J_I| > 1");
    const Irrep& ir = P.bI->irrep[gammaI];
    GAIN_CHECK_ARG(idegI >= 0 && idegI < ir.sdeg, "component out of range");
    int slot, cslot;
    char* hs = (char*)c->ring.acquire(sizeof(int64_t) + sizeof(int), &slot);
    *(int64_t*)hs = 0;
    *(int*)(hs + sizeof(int64_t)) = idegI;
    char* ds = c->w_small.get<char>(sizeof(int64_t) + sizeof(int));
    GAIN_HIP(hipMemcpyAsync(ds, hs, sizeof(int64_t) + sizeof(int),
                            hipMemcpyHostToDevice, c->stream));
    const size_t nb = sizeof(double) * std::max(ir.ntot, 1);
    void* hc = c->ring.acquire(nb, &cslot);
    std::memcpy(hc, cI, sizeof(double) * ir.ntot);
    double* dc = c->w_c.get<double>(std::max(ir.ntot, 1));
    GAIN_HIP(hipMemcpyAsync(dc, hc, nb, hipMemcpyHostToDevice, c->stream));
    c->ring.release(slot, c->stream);
    c->ring.release(cslot, c->stream);
    Small sm{};
    sm.c_off = (int64_t*)ds;
    sm.ideg = (int*)(ds + sizeof(int64_t));
    double* H = c->w_H.get<double>(P.bF->nprim);
    run_half_ls(c, P, dc, sm, 1, H, GAIN_PREC_FP64);
    GAIN_HIP(hipMemcpyAsync(h_out, H, sizeof(double) * P.bF->nprim,
                            hipMemcpyDeviceToHost, c->stream));
    GAIN_HIP(hipStreamSynchronize(c->stream));
  });
}

extern "C" int gain_plain_lines(gain_ctx* c, int jindI, int gammaI,
                                int levelI, const double* cI, int jindF,
                                int gammaF, const gain_lines* out) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(out, "null output descriptor");
    DeviceGuard g(c->device);
    Plan P = make_plan(c, jindI, gammaI, &levelI, 1, jindF, gammaF);
    run_lines(c, P, cI, GAIN_PREC_FP64, true, out);
  });
}

extern "C" int gain_batched_lines(gain_ctx* c, int jindI, int gammaI,
                                  const int* levelsI, int nI, int jindF,
                                  int gammaF, int precision,
                                  const gain_lines* out) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(out && levelsI && nI > 0, "bad batch");
    GAIN_CHECK_ARG(precision >= GAIN_PREC_FP64 &&
                       precision <= GAIN_PREC_I8CRT14,
                   "unknown precision %d", precision);
    DeviceGuard g(c->device);
    Plan P = make_plan(c, jindI, gammaI, levelsI, nI, jindF, gammaF);
    if (!P.blkI->vec)
      fail(GAIN_ERR_STATE, "batched path needs the initial block's vectors "
                           "resident (gain_upload_block with vectors)");
    run_lines(c, P, nullptr, precision, false, out);
  });
}

extern "C" int gain_set_timing(gain_ctx* c, int enable) {
  return guarded([&] {
    need(c);
    c->timing = enable != 0;
  });
}

extern "C" int gain_get_timing(gain_ctx* c, const char* name, double* total_ms,
                               int64_t* launches, int reset) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(name && total_ms && launches, "null argument");
    DeviceGuard g(c->device);
    auto it = c->timers.find(name);
    if (it == c->timers.end()) {
      *total_ms = 0.0;
      *launches = 0;
      return;
    }
    resolve_timers(c, it->second);
    *total_ms = it->second.total_ms;
    *launches = it->second.launches;
    if (reset) {
      it->second.total_ms = 0.0;
      it->second.launches = 0;
    }
  });
}

// Tuning knobs (not part of the public header; used by the autotuner).
extern "C" int gain_set_tuning(gain_ctx* c, const char* key, int value) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(key, "null key");
    const std::string k(key);
    if (k == "gemv_rows") {
      GAIN_CHECK_ARG(value == 2 || value == 4 || value == 8, "rows 2/4/8");
      c->gemv_rows = value;
    } else if (k == "gemv_ksplit") {
      GAIN_CHECK_ARG(value >= 1 && value <= 256, "ksplit 1..256");
      c->gemv_ksplit = value;
    } else if (k == "mu_ksplit") {
      GAIN_CHECK_ARG(value >= 0 && value <= 64, "ksplit 0..64");
      c->mu_ksplit = value;
    } else if (k == "mfma") {
      c->use_mfma = value != 0;
    } else if (k == "autotune") {
      c->autotune = value != 0;
      if (value == 2) c->tuned.clear();
    } else if (k == "proj_split") {
      GAIN_CHECK_ARG(value >= 1 && value <= 64, "split 1..64");
      c->force_split = value;
    } else if (k == "i8_cfg") {
      GAIN_CHECK_ARG(value >= -1 && value < kI8Configs, "cfg -1..%d",
                     kI8Configs - 1);
      c->force_i8_cfg = value;
    } else if (k == "crt_cfg") {
      GAIN_CHECK_ARG(value >= -1 && value < kCrtConfigs, "cfg -1..%d",
                     kCrtConfigs - 1);
      c->force_crt_cfg = value;
    } else if (k == "crt_kskip") {
      c->crt_kskip = value != 0;
    } else if (k == "crt_fuse") {
      c->crt_fuse = value != 0;
    } else if (k == "crt_xfuse") {
      c->crt_xfuse = value != 0;
    } else if (k == "f32_cfg") {
      GAIN_CHECK_ARG(value >= -1 && value < kF32Configs, "cfg -1..%d",
                     kF32Configs - 1);
      c->f32_cfg = value;
    } else if (k == "proj_crt_moduli") {
      GAIN_CHECK_ARG(value == 0 || value == 11 || value == 12 || value == 14,
                     "projection moduli 0 (as the precision), 11, 12 or 14");
      c->proj_crt = value;
    } else if (k == "i8_split") {
      GAIN_CHECK_ARG(value >= 0 && value <= 4, "split 0..4");
      c->force_i8_split = value;
    } else if (k == "sweep_phases") {
      GAIN_CHECK_ARG(value >= 1 && value <= 16, "phases 1..16");
      c->sweep_phases = value;
    } else if (k == "sweep_eager") {
      c->sweep_eager = value != 0;
    } else if (k == "noalloc_cap") {
      GAIN_CHECK_ARG(value >= 0, "cap >= 0 bytes (0: the whole array)");
      c->noalloc_cap = value;
    } else if (k == "proj_cfg" || k == "half_ls_cfg") {
      GAIN_CHECK_ARG(value >= -1 && value < kGemmConfigs, "cfg -1..%d",
                     kGemmConfigs - 1);
      c->force_cfg[k == "proj_cfg" ? 0 : 1] = value;
    } else {
      fail(GAIN_ERR_ARG, "unknown tuning key %s", key);
    }
  });
}

// dL/dmu of one batch (both entry points): the FP64 forward recomputed,
// or the forward's pair line strengths and flags given (ls_in, flags_in,
// device or host memory as dS).
static int adjoint_entry(gain_ctx* c, int jindI, int gammaI,
                         const int* levelsI, int nI, int jindF, int gammaF,
                         const double* dS, bool given, const double* ls_in,
                         const uint8_t* flags_in, double* dmu, int on_device,
                         int accumulate) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(levelsI && nI > 0 && dS && dmu, "bad adjoint arguments");
    GAIN_CHECK_ARG(!given || (ls_in && flags_in),
                   "the forward's line strengths and flags are required");
    DeviceGuard g(c->device);
    Plan P = make_plan(c, jindI, gammaI, levelsI, nI, jindF, gammaF);
    if (!P.blkI->vec)
      fail(GAIN_ERR_STATE, "adjoint needs the initial block's vectors "
                           "resident");
    const int nc = c->nc;
    const size_t nF = P.blkF->n;
    const size_t nmu = (size_t)3 * nc * nc;
    const size_t md2 = (size_t)c->maxdeg * c->maxdeg;
    const double* dS_dev = dS;
    const double* ls_dev = ls_in;
    const uint8_t* fl_dev = flags_in;
    if (!on_device) {
      double* d = c->w_c.get<double>((size_t)nI * nF);
      GAIN_HIP(hipMemcpyAsync(d, dS, sizeof(double) * nI * nF,
                              hipMemcpyHostToDevice, c->stream));
      dS_dev = d;
      if (ls_in) {
        double* l = c->w_gs.get<double>((size_t)nI * nF * md2);
        GAIN_HIP(hipMemcpyAsync(l, ls_in, sizeof(double) * nI * nF * md2,
                                hipMemcpyHostToDevice, c->stream));
        ls_dev = l;
        uint8_t* f = c->w_ge.get<uint8_t>((size_t)nI * nF);
        GAIN_HIP(hipMemcpyAsync(f, flags_in, (size_t)nI * nF,
                                hipMemcpyHostToDevice, c->stream));
        fl_dev = f;
      }
    }
    double* work = c->w_xs.get<double>(nmu);
    run_adjoint(c, P, dS_dev, work, ls_dev, fl_dev);
    if (on_device) {
      launch_axpy1(work, dmu, (int64_t)nmu, accumulate, c->stream);
    } else {
      std::vector<double> h(nmu);
      GAIN_HIP(hipMemcpyAsync(h.data(), work, sizeof(double) * nmu,
                              hipMemcpyDeviceToHost, c->stream));
      GAIN_HIP(hipStreamSynchronize(c->stream));
      for (size_t i = 0; i < nmu; ++i)
        dmu[i] = accumulate ? dmu[i] + h[i] : h[i];
    }
    GAIN_HIP(hipStreamSynchronize(c->stream));
  });
}

extern "C" int gain_adjoint_dipole(gain_ctx* c, int jindI, int gammaI,
                                   const int* levelsI, int nI, int jindF,
                                   int gammaF, const double* dS, double* dmu,
                                   int on_device, int accumulate) {
  return adjoint_entry(c, jindI, gammaI, levelsI, nI, jindF, gammaF, dS,
                       false, nullptr, nullptr, dmu, on_device, accumulate);
}

extern "C" int gain_adjoint_dipole_ls(gain_ctx* c, int jindI, int gammaI,
                                      const int* levelsI, int nI, int jindF,
                                      int gammaF, const double* dS,
                                      const double* ls, const uint8_t* flags,
                                      double* dmu, int on_device,
                                      int accumulate) {
  return adjoint_entry(c, jindI, gammaI, levelsI, nI, jindF, gammaF, dS,
                       true, ls, flags, dmu, on_device, accumulate);
}

extern "C" int gain_set_dipole_combination(gain_ctx* c, int P,
                                           const double* coef,
                                           const double* M_dev, int nc) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(P > 0 && coef && M_dev && nc > 0, "bad combination");
    DeviceGuard g(c->device);
    if (c->d_mu && c->nc != nc)
      fail(GAIN_ERR_ARG, "dipole dimension %d differs from %d", nc, c->nc);
    ++c->gen;
    release_sliced(c, c->mu_sl);
    if (!c->d_mu) {
      c->d_mu = c->arena.alloc<double>((size_t)3 * nc * nc);
      c->nc = nc;
      build_slots_all(c);
    }
    c->mu_have.assign(nc, 1);  // the combination fills every column
    c->mu_missing = 0;
    double* dcoef = c->w_c.get<double>(P);
    GAIN_HIP(hipMemcpyAsync(dcoef, coef, sizeof(double) * P,
                            hipMemcpyHostToDevice, c->stream));
    launch_combine(P, (int64_t)3 * nc * nc, dcoef, M_dev, c->d_mu,
                   c->stream);
    GAIN_HIP(hipStreamSynchronize(c->stream));
  });
}

extern "C" int gain_dipole_coeff_grad(gain_ctx* c, int P, const double* M_dev,
                                      const double* dmu_dev, double* dc) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(P > 0 && M_dev && dmu_dev && dc, "bad arguments");
    DeviceGuard g(c->device);
    const int64_t n = (int64_t)3 * c->nc * c->nc;
    GAIN_CHECK_ARG(n > 0 && n < (int64_t)1 << 31, "dipole too large");
    const int ks = 256;
    double* part = c->w_part.get<double>((size_t)ks * P);
    GemvArgs a{};
    a.C = M_dev;
    a.ld = n;
    a.ntot = (int)n;
    a.F0 = 0;
    a.nF = P;
    a.G = dmu_dev;
    a.gstride = n;
    a.nvec = 1;
    a.partial = part;
    a.ksplit = ks;
    a.rows = 4;
    launch_gemv(a, c->stream);
    std::vector<double> h((size_t)ks * P);
    GAIN_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * ks * P,
                            hipMemcpyDeviceToHost, c->stream));
    GAIN_HIP(hipStreamSynchronize(c->stream));
    for (int q = 0; q < P; ++q) {
      double s = 0.0;
      for (int k = 0; k < ks; ++k) s += h[(size_t)k * P + q];
      dc[q] = s;
    }
  });
}

extern "C" int gain_get_tuning(gain_ctx* c, const char* key,
                               long long* value) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(key && value, "null argument");
    const std::string k(key);
    if (k == "gemv_rows") {
      *value = c->gemv_rows;
    } else if (k == "gemv_ksplit") {
      *value = c->gemv_ksplit;
    } else if (k == "mu_ksplit") {
      *value = c->mu_ksplit;
    } else if (k == "mfma") {
      *value = c->use_mfma;
    } else if (k == "autotune") {
      *value = c->autotune;
    } else if (k == "crt_kskip") {
      *value = c->crt_kskip;
    } else if (k == "crt_fuse") {
      *value = c->crt_fuse;
    } else if (k == "crt_xfuse") {
      *value = c->crt_xfuse;
    } else if (k == "proj_crt_moduli") {
      *value = c->proj_crt;
    } else if (k == "f32_cfg") {
      *value = c->f32_cfg;
    } else if (k == "autotune_us") {
      auto it = c->last.find(k);  // nothing tuned yet: 0
      *value = it == c->last.end() ? 0 : it->second;
    } else {
      auto it = c->last.find(k);
      if (it == c->last.end())
        fail(GAIN_ERR_ARG, "unknown tuning key %s, or no launch yet", key);
      *value = it->second;
    }
  });
}

extern "C" int gain_tuning_report(gain_ctx* c, char* buf, int len) {
  return guarded([&] {
    need(c);
    GAIN_CHECK_ARG(buf && len > 0, "bad buffer");
    snprintf(buf, len, "%s", c->tune_log.c_str());
  });
}

extern "C" int gain_crt_moduli(int nm, int* m) {
  return guarded([&] {
    GAIN_CHECK_ARG(m, "null argument");
    for (int i = 0; i < nm; ++i) m[i] = crt_modulus(nm, i);
  });
}

extern "C" int gain_crt_beta(int nm, long long K, int* beta,
                             int* norm_scaled) {
  return guarded([&] {
    GAIN_CHECK_ARG(beta && K > 0 && (nm == 11 || nm == 12 || nm == 14),
                   "bad CRT width arguments");
    *beta = crt_beta(nm, K);
    if (norm_scaled) *norm_scaled = crt_norm_scaled(nm, K) ? 1 : 0;
  });
}

extern "C" int gain_crt_reconstruct(int nm, const unsigned int* words,
                                    long long n4, double* out) {
  return guarded([&] {
    GAIN_CHECK_ARG(words && out && n4 >= 0, "bad reconstruction arguments");
    crt_reconstruct_host(nm, words, n4, out);
  });
}
