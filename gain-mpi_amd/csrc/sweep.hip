// Mixed-J sweeps: many (initial batch -> final block) tasks in one launch
// sequence captured as a HIP graph.
//
// Tasks with the same initial batch (J index, irrep, levels) share X and
// Y = mu X.  The X of every batch sits in one column range of a single
// matrix; the batches are ordered by J_I and cut into phases of about equal
// column counts, and mu X of a phase is one GEMM over its columns (tile
// shape from the autotuner; FP64, FP32 products, or split-BF16, split-INT8
// or CRT form on the sweep's own copy of mu).  The products run phase after
// phase on the context's stream; everything downstream of a phase
// (gathers, transposed
// expansions, projections, epilogues) runs on a second stream once that
// phase's product is done, so the HBM-bound projections of finished phases
// run beside the matrix-core work of the next phase.  A final block of J_F
// receives tasks from J_I = J_F - 1, J_F, J_F + 1 only, so with J-ordered
// phases its projection is ready soon after its initial batches.
// The projections C_F G of the tasks with many vectors run as one
// persistent grouped launch per phase over the tiles of all of them.
// Tasks that project at most 9 vectors go through the FP64 GEMV instead (a
// 128-wide GEMM tile would be mostly padding); such tasks that target the
// same final block are packed into one GEMV of up to 9 vectors over the
// union of their energy windows, so the block's records are read once per
// pack rather than once per task; a pack runs in the phase of its last
// task, and the packs of one phase run in one grouped GEMV launch.  The
// per-batch and per-task kernels (scatter, gather, transposed expansion,
// epilogue) also run as one grouped launch per stage and phase over
// argument tables built at creation.
// Everything a replay touches is owned by the sweep (sized at creation), so
// replays stay valid whatever the context does in between.  The launch
// sequence contains kernels and event record/wait pairs only (no memset or
// memcpy nodes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

#include "engine_internal.hpp"

using namespace gain;

struct gain_sweep {
  gain_ctx* c = nullptr;
  DeviceArena arena;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t side = nullptr;  // downstream stages of every phase
  hipEvent_t join = nullptr;   // side stream done (recorded once per run)
  struct Group {
    Plan P;
    int nvecX = 0, ncolTot = 0;
    int phase = 0;
    double* X = nullptr;
    double* Y = nullptr;
    int64_t* c_off = nullptr;
    int* ideg = nullptr;
  };
  struct Task {
    Plan P;
    int group = 0;
    int phase = 0;  // phase of its initial batch
    int ready = 0;  // phase whose side stage projects it
    int nvpi = 0, nvecG = 0, nR = 0;
    int64_t gstride = 0, hstride = 0;
    bool compute = false;
    bool gemv = false;  // projection by GEMV (few vectors), else GEMM tiles
    int unit = -1, vecoff = 0;  // GEMV pack and first vector inside it
    double* H = nullptr;
    double* G = nullptr;
    double* LS = nullptr;
    int* h_of_vec = nullptr;
    int* ideg_of_vec = nullptr;
    int* slot_of_vec = nullptr;
    double* eI = nullptr;
    int* levelI = nullptr;
    uint8_t* okI = nullptr;
    gain_lines out{};
  };
  // one GEMV over one final block for a pack of tasks
  struct Unit {
    const Block* blkF = nullptr;
    int ntot = 0;
    int64_t gstride = 0;
    int Flo = 0, Fhi = 0, nvec = 0, ksplit = 1;
    int ready = 0;  // latest phase among its tasks
    double* G = nullptr;
    double* part = nullptr;  // [ksplit][Fhi-Flo][nvec]
  };
  // one grouped launch per stage: argument table + workgroup offsets
  struct Launch {
    int kind = 0;
    const void* args = nullptr;
    const int* start = nullptr;
    int n = 0, blocks = 0;
  };
  struct Phase {
    int64_t col0 = 0, ncol = 0;  // column range of X and Y (col0 % 256 == 0)
    bool has_hl = false;
    GemmArgs hl{};    // FP64 (or FP32-product) mu X of the phase's columns
    I8GemmArgs hl8{};  // split-INT8 / CRT form of the same
    SlicedGemmArgs hlb{};  // split-BF16 form: X of the phase cut into
    uint16_t* xs = nullptr;  // its own bf16 planes [S][ncol][Kp]
    int* xe = nullptr;       // and chunk exponents [ncol][nch]
    std::vector<Launch> scatter, gather, expand, gemv, lines;
    std::vector<GatherRotArgs> gather_rot;  // symmetric-top tasks, one
                                            // launch each
    std::vector<GemvArgs> gemv_single;  // units whose layout the grouped
                                        // GEMV cannot take
    GemmArgs* d_pj_probs = nullptr;
    GroupTile* d_pj_tiles = nullptr;
    int n_pj_tiles = 0;
    hipEvent_t done = nullptr;  // the phase's mu X is complete
  };
  std::vector<Group> groups;
  std::vector<Task> tasks;
  std::vector<Unit> units;
  std::vector<Phase> phases;
  int64_t ncol_all = 0;
  double* X = nullptr;
  double* Y = nullptr;
  // split-INT8 (GAIN_PREC_I8X*) or CRT-form (GAIN_PREC_I8CRT*) mu X: the
  // sweep's own int8 copy of mu, X cut into planes at every replay
  int slices = 0;
  int crt = 0;  // moduli of the CRT form, 0 otherwise
  // FP32-product (GAIN_PREC_FP32) or split-BF16 (GAIN_PREC_BF16X*, the
  // sweep's own bf16 slices of mu) mu X
  bool f32 = false;
  int f32_cfg = -1;  // its tile configuration (the context's at creation)
  int bf16 = 0;
  int8_t* x8 = nullptr;
  int* x8e = nullptr;
  int64_t x8_plane = 0;
  int pj_cfg = 5;
  int grid = 1024;
  uint64_t gen = 0;  // context generation the captured pointers belong to
};

namespace {

template <class T>
T* up(gain_sweep* s, const std::vector<T>& v) {
  return s->arena.upload(v.data(), v.size());
}

void add_tiles(std::vector<GroupTile>& tiles, int prob, int M, int N, int nz,
               int cfg) {
  const int bm = gemm_tile_m(cfg), bn = gemm_tile_n(cfg);
  const int ntm = (M + bm - 1) / bm, ntn = (N + bn - 1) / bn;
  for (int z = 0; z < nz; ++z)
    for (int tm = 0; tm < ntm; ++tm)
      for (int tn = 0; tn < ntn; ++tn) tiles.push_back({prob, tm, tn, z});
}

template <class A>
void add_stage(gain_sweep* s, std::vector<gain_sweep::Launch>& out, int kind,
               const std::vector<A>& v, const GroupedTable& tab) {
  if (tab.keep.empty()) return;
  std::vector<A> kept;
  for (int t : tab.keep) kept.push_back(v[t]);
  gain_sweep::Launch L;
  L.kind = kind;
  L.args = up(s, kept);
  L.start = up(s, tab.start);
  L.n = (int)tab.keep.size();
  L.blocks = tab.start.back();
  out.push_back(L);
}

void run_stage(const std::vector<gain_sweep::Launch>& v, hipStream_t st) {
  for (const gain_sweep::Launch& L : v)
    launch_grouped(L.kind, L.args, L.start, L.n, L.blocks, st);
}

// Phases: the batches ordered by J_I (then by creation order) and cut where
// the running column count passes p/P of the total (P from the tuning key
// "sweep_phases", at most the number of batches); each phase's columns
// start on a multiple of 256, so the int8 GEMM of a phase reads no plane
// row of the next one.  Returns the batches of each phase.
std::vector<std::vector<int>> assign_phases(gain_sweep* s) {
  std::vector<int> order;
  int64_t total = 0;
  for (size_t g = 0; g < s->groups.size(); ++g)
    if (s->groups[g].nvecX > 0) {
      order.push_back((int)g);
      total += s->groups[g].ncolTot;
    }
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return s->groups[x].P.bI->J < s->groups[y].P.bI->J;
  });
  const int P =
      std::max(1, std::min<int>(s->c->sweep_phases, (int)order.size()));
  std::vector<std::vector<int>> cut(P);
  int64_t run = 0;
  int p = 0;
  for (int g : order) {
    cut[p].push_back(g);
    run += s->groups[g].ncolTot;
    if (p < P - 1 && run * P >= total * (p + 1)) ++p;
  }
  std::vector<std::vector<int>> members;
  for (auto& v : cut)
    if (!v.empty()) members.push_back(v);
  if (members.empty()) members.emplace_back();  // a sweep without products
  s->phases.assign(members.size(), gain_sweep::Phase{});
  int64_t col = 0;
  for (size_t pi = 0; pi < members.size(); ++pi) {
    col = (col + 255) / 256 * 256;
    s->phases[pi].col0 = col;
    for (int g : members[pi]) {
      s->groups[g].phase = (int)pi;
      col += s->groups[g].ncolTot;
    }
    s->phases[pi].ncol = col - s->phases[pi].col0;
  }
  s->ncol_all = s->phases.back().col0 + s->phases.back().ncol;
  return members;
}

// Scatter launches of every phase (built before the products are tuned, so
// that the tuner times the real X).
void build_scatter(gain_sweep* s, const std::vector<std::vector<int>>& mem) {
  gain_ctx* c = s->c;
  const int nc = c->nc;
  for (size_t pi = 0; pi < mem.size(); ++pi) {
    std::vector<ScatterArgs> sc;
    for (int g : mem[pi]) {
      gain_sweep::Group& G = s->groups[g];
      Basis& bI = *G.P.bI;
      const Irrep& ir = bI.irrep[G.P.gI];
      ScatterArgs sa{};
      sa.nslot = bI.ncol * nc;
      sa.nprim = bI.nprim;
      sa.slot_ptr = bI.slot_ptr;
      sa.slot_prim = bI.slot_prim;
      sa.kval = bI.kval;
      sa.tau = bI.tau;
      sa.icontr = bI.icontr;
      sa.ijt = bI.ijterms + (int64_t)G.P.gI * bI.maxsymcoeffs;
      sa.N = ir.N;
      sa.repres = ir.repres;
      sa.sdeg = ir.sdeg;
      sa.ntot = ir.ntot;
      sa.cbase = G.P.blkI->vec;
      sa.c_off = G.c_off;
      sa.ideg = G.ideg;
      sa.nvec = G.nvecX;
      sa.X = G.X;
      sa.xstride = (int64_t)bI.ncol * nc;
      sc.push_back(sa);
    }
    add_stage(s, s->phases[pi].scatter, kGroupedScatter, sc,
              grouped_scatter_table(sc));
  }
}

// mu X of one phase on stream st: the FP64 GEMM, or X cut into int8 planes
// and the split-INT8 / CRT-form products.
void products(gain_sweep* s, const gain_sweep::Phase& ph, hipStream_t st) {
  if (!ph.has_hl) return;
  const int nc = s->c->nc;
  const double* X = s->X + ph.col0 * nc;
  if (s->f32) {
    gemm_f32acc(ph.hl, st, s->f32_cfg);
    return;
  }
  if (s->bf16 > 0) {
    slice_rows_bf16(s->bf16, X, nc, (int)ph.ncol, nc, ph.xs, ph.xe, st);
    gemm_sliced_bf16(ph.hlb, st);
    return;
  }
  if (s->slices == 0 && s->crt == 0) {
    gemm_f64(ph.hl, st);
    return;
  }
  int8_t* planes = (int8_t*)ph.hl8.B;
  if (s->crt > 0) {
    slice_rows_crt(s->crt, ph.hl8.beta, X, nc, (int)ph.ncol, nc, planes,
                   s->x8_plane, (int*)ph.hl8.eB, st);
    gemm_crt_i8(ph.hl8, st);
  } else {
    slice_rows_i8(s->slices, X, nc, (int)ph.ncol, nc, planes, s->x8_plane,
                  (int*)ph.hl8.eB, st);
    gemm_sliced_i8(ph.hl8, st);
  }
}

void build_launches(gain_sweep* s);

void build(gain_sweep* s, const gain_task* tasks, int ntasks) {
  gain_ctx* c = s->c;
  const int nc = c->nc;
  std::map<std::tuple<int, int, std::vector<int>>, int> gidx;
  for (int t = 0; t < ntasks; ++t) {
    const gain_task& tk = tasks[t];
    GAIN_CHECK_ARG(tk.levelsI && tk.nI > 0, "task %d: no initial states", t);
    GAIN_CHECK_ARG(tk.out.on_device, "task %d: outputs must be device "
                   "memory (replays write them)", t);
    gain_sweep::Task T;
    T.P = make_plan(c, tk.jindI, tk.gammaI, tk.levelsI, tk.nI, tk.jindF,
                    tk.gammaF);
    if (!T.P.blkI->vec)
      fail(GAIN_ERR_STATE, "task %d: initial block has no resident vectors",
           t);
    T.out = tk.out;
    const auto key = std::make_tuple(tk.jindI, tk.gammaI, T.P.levels);
    auto it = gidx.find(key);
    if (it == gidx.end()) {
      gain_sweep::Group G;
      G.P = T.P;
      gidx[key] = (int)s->groups.size();
      s->groups.push_back(G);
      it = gidx.find(key);
    }
    T.group = it->second;
    // the d_I set depends on the final irrep only through the pair rule,
    // which is the same for every final block of a valid task
    if (T.P.DI != s->groups[T.group].P.DI)
      fail(GAIN_ERR_ARG, "task %d: initial components differ within a batch",
           t);
    s->tasks.push_back(T);
  }
  // ---- groups: scatter arrays, phases, X and Y (column ranges of one
  // matrix)
  for (size_t g = 0; g < s->groups.size(); ++g) {
    gain_sweep::Group& G = s->groups[g];
    Basis& bI = *G.P.bI;
    ensure_slots(c, bI);
    const int nDI = (int)G.P.DI.size();
    G.nvecX = G.P.nI * nDI;
    G.ncolTot = G.nvecX * bI.ncol;
    std::vector<int64_t> off;
    std::vector<int> deg;
    for (int i = 0; i < G.P.nI; ++i)
      for (int k = 0; k < nDI; ++k) {
        off.push_back((int64_t)G.P.levels[i] * G.P.blkI->ld);
        deg.push_back(G.P.DI[k]);
      }
    if (G.nvecX == 0) continue;
    G.c_off = up(s, off);
    G.ideg = up(s, deg);
  }
  const std::vector<std::vector<int>> mem = assign_phases(s);
  for (gain_sweep::Phase& ph : s->phases)
    GAIN_HIP(hipEventCreateWithFlags(&ph.done, hipEventDisableTiming));
  if (s->ncol_all > 0) {
    GAIN_CHECK_ARG(s->ncol_all < (1ll << 31), "sweep: too many columns");
    s->X = s->arena.alloc<double>((size_t)s->ncol_all * nc);
    s->Y = s->arena.alloc<double>((size_t)3 * s->ncol_all * nc);
    // the columns between phases stay zero (cleared once, outside the graph)
    GAIN_HIP(hipMemsetAsync(s->X, 0, sizeof(double) * s->ncol_all * nc,
                            c->stream));
    for (size_t pi = 0; pi < mem.size(); ++pi) {
      int64_t col = s->phases[pi].col0;
      for (int g : mem[pi]) {
        gain_sweep::Group& G = s->groups[g];
        G.X = s->X + col * nc;
        G.Y = s->Y + col * nc;  // alpha planes are ncol_all * nc apart
        col += G.ncolTot;
      }
    }
    build_scatter(s, mem);
    for (const gain_sweep::Phase& ph : s->phases)
      run_stage(ph.scatter, c->stream);
    // int8 copy of mu (split-INT8 or CRT form)
    int8_t* m8 = nullptr;
    int* me = nullptr;
    int2* mkr = nullptr;  // CRT form: nonzero k blocks of mu's row blocks
    int beta = 0;
    const int64_t Kp = i8_kpad(nc), Rm = i8_rows_pad(nc);
    if (s->crt > 0 || s->slices > 0) {
      const int S = s->crt > 0 ? s->crt : s->slices;
      double* muT = s->arena.alloc<double>((size_t)3 * nc * nc);
      transpose_mu(c->d_mu, nc, muT, c->stream);
      m8 = s->arena.alloc<int8_t>((size_t)S * 3 * Rm * Kp);
      me = s->arena.alloc<int>((size_t)3 * Rm);
      if (s->crt > 0) beta = crt_beta(s->crt, Kp);
      for (int al = 0; al < 3; ++al) {
        if (s->crt > 0)
          slice_rows_crt(s->crt, beta, muT + (size_t)al * nc * nc, nc, nc, nc,
                         m8 + al * Rm * Kp, 3 * Rm * Kp, me + al * Rm,
                         c->stream);
        else
          slice_rows_i8(S, muT + (size_t)al * nc * nc, nc, nc, nc,
                        m8 + al * Rm * Kp, 3 * Rm * Kp, me + al * Rm,
                        c->stream);
      }
      if (s->crt > 0) {
        mkr = s->arena.alloc<int2>((size_t)3 * (Rm / 32));
        i8_block_kranges(m8, 3 * Rm * Kp, s->crt, Rm * Kp, 3, (int)(Rm / 32),
                         (int)(Kp / 32), mkr, c->stream);
      }
      GAIN_HIP(hipStreamSynchronize(c->stream));
      s->arena.release(muT);
      const int64_t Rx = i8_rows_pad(s->ncol_all);
      s->x8_plane = Rx * Kp;
      s->x8 = s->arena.alloc<int8_t>((size_t)S * s->x8_plane);
      s->x8e = s->arena.alloc<int>((size_t)Rx);
    }
    // split-BF16: the sweep's own bf16 slices of mu along v_I
    uint16_t* mb = nullptr;
    int* mbe = nullptr;
    const int64_t Kpb = slice_kpad(nc);
    const int nchb = s->bf16 > 0 ? (int)slice_nchunk(s->bf16, nc) : 0;
    if (s->bf16 > 0) {
      double* muT = s->arena.alloc<double>((size_t)3 * nc * nc);
      transpose_mu(c->d_mu, nc, muT, c->stream);
      mb = s->arena.alloc<uint16_t>((size_t)s->bf16 * 3 * nc * Kpb);
      mbe = s->arena.alloc<int>((size_t)3 * nc * nchb);
      slice_rows_bf16(s->bf16, muT, nc, 3 * nc, nc, mb, mbe, c->stream);
      GAIN_HIP(hipStreamSynchronize(c->stream));
      s->arena.release(muT);
    }
    int64_t rb = 0;  // CRT residue buffer for every phase and configuration
    for (gain_sweep::Phase& ph : s->phases) {
      if (ph.ncol == 0) continue;
      ph.has_hl = true;
      GemmArgs& a = ph.hl;
      a.M = nc;
      a.N = (int)ph.ncol;
      a.K = nc;
      a.A = c->d_mu;
      a.lda = nc;
      a.sAz = (int64_t)nc * nc;
      a.a_kcontig = false;
      a.B = s->X + ph.col0 * nc;
      a.ldb = nc;
      a.C = s->Y + ph.col0 * nc;
      a.ldc = nc;
      a.sCz = s->ncol_all * nc;
      a.nz = 3;
      a.ksplit = 1;
      if (s->f32) continue;  // the context's FP32 tile configuration
      if (s->bf16 > 0) {
        ph.xs = s->arena.alloc<uint16_t>((size_t)s->bf16 * ph.ncol * Kpb);
        ph.xe = s->arena.alloc<int>((size_t)ph.ncol * nchb);
        SlicedGemmArgs& g = ph.hlb;
        g.S = s->bf16;
        g.M = nc;
        g.N = (int)ph.ncol;
        g.Kp = Kpb;
        g.A = mb;
        g.a_plane = (int64_t)3 * nc * Kpb;
        g.a_z = (int64_t)nc * Kpb;
        g.eA = mbe;
        g.ea_z = (int64_t)nc * nchb;
        g.nch = nchb;
        g.B = ph.xs;
        g.b_plane = ph.ncol * Kpb;
        g.eB = ph.xe;
        g.C = s->Y + ph.col0 * nc;
        g.ldc = nc;
        g.c_z = s->ncol_all * nc;
        g.nz = 3;
        continue;
      }
      if (s->crt == 0 && s->slices == 0) {
        a.cfg = gemm_cfg(c, 1, a);  // tuned here: replays must not time
        continue;
      }
      I8GemmArgs& g = ph.hl8;
      g.S = s->crt > 0 ? s->crt : s->slices;
      g.beta = beta;
      g.M = nc;
      g.N = (int)ph.ncol;
      g.row0 = 0;
      g.Kp = Kp;
      g.A = m8;
      g.a_plane = 3 * Rm * Kp;
      g.a_z = Rm * Kp;
      g.eA = me;
      g.ea_z = Rm;
      g.B = s->x8 + ph.col0 * Kp;  // row blocks of 32 rows x Kp bytes
      g.b_plane = s->x8_plane;
      g.eB = s->x8e + ph.col0;
      g.C = s->Y + ph.col0 * nc;
      g.ldc = nc;
      g.c_z = s->ncol_all * nc;
      g.c_ks = 3 * s->ncol_all * nc;
      g.nz = 3;
      if (s->crt > 0 && c->crt_kskip) {
        g.a_kr = mkr;
        g.akr_z = Rm / 32;
        g.akr_n = (int)(Rm / 32);
      }
      if (s->crt > 0)
        for (int cfg = 0; cfg < kCrtConfigs; ++cfg) {
          g.cfg = cfg;
          rb = std::max(rb, gemm_crt_res_bytes(g));
        }
    }
    int8_t* R = rb > 0 ? s->arena.alloc<int8_t>((size_t)rb) : nullptr;
    for (gain_sweep::Phase& ph : s->phases) {
      if (!ph.has_hl || (s->crt == 0 && s->slices == 0)) continue;
      I8GemmArgs& g = ph.hl8;
      const double* X = s->X + ph.col0 * nc;
      // cut the phase's X so the tuner times real operands
      if (s->crt > 0) {
        g.R = R;
        slice_rows_crt(s->crt, beta, X, nc, (int)ph.ncol, nc, (int8_t*)g.B,
                       s->x8_plane, (int*)g.eB, c->stream);
        g.cfg = crt_cfg_of(c, 1, g);
      } else {
        slice_rows_i8(s->slices, X, nc, (int)ph.ncol, nc, (int8_t*)g.B,
                      s->x8_plane, (int*)g.eB, c->stream);
        int ks = 1;
        g.cfg = i8_cfg_split(c, 1, g, 1, &ks);
        g.ksplit = ks;
      }
    }
  }
  // ---- tasks: H, G, LS, per-pair arrays, projection problems per phase
  std::vector<std::vector<GemmArgs>> pj(s->phases.size());
  std::vector<std::vector<GroupTile>> pjt(s->phases.size());
  for (gain_sweep::Task& T : s->tasks) {
    const gain_sweep::Group& G = s->groups[T.group];
    T.phase = T.ready = G.phase;
    Basis& bF = *T.P.bF;
    const Irrep& irF = bF.irrep[T.P.gF];
    const int nDI = (int)T.P.DI.size();
    T.nvpi = (int)T.P.vec.size();
    T.nvecG = T.P.nI * T.nvpi;
    T.nR = T.P.Fhi - T.P.Flo;
    T.compute = T.nR > 0 && T.nvpi > 0 && nDI > 0;
    T.gstride = (irF.ntot + 1) / 2 * 2;
    T.hstride = bF.nprim;
    std::vector<int> hv, dv, sv;
    for (int i = 0; i < T.P.nI; ++i)
      for (int v = 0; v < T.nvpi; ++v) {
        const int dI = T.P.vec[v].first;
        int pos = -1;
        for (int k = 0; k < nDI; ++k)
          if (T.P.DI[k] == dI) pos = k;
        hv.push_back(pos >= 0 ? i * nDI + pos : G.nvecX);
        dv.push_back(T.P.vec[v].second);
      }
    for (int v = 0; v < T.nvpi; ++v)
      sv.push_back(T.P.vec[v].first + T.P.vec[v].second * c->maxdeg);
    sv.push_back(0);
    T.slot_of_vec = up(s, sv);
    T.eI = up(s, T.P.eI);
    T.levelI = up(s, T.P.levels);
    T.okI = up(s, T.P.okI);
    if (!T.compute) continue;
    T.h_of_vec = up(s, hv);
    T.ideg_of_vec = up(s, dv);
    T.H = s->arena.alloc<double>((size_t)(G.nvecX + 1) * T.hstride);
    // the trailing zero vector is constant: cleared once, outside the graph
    GAIN_HIP(hipMemset(T.H + (size_t)G.nvecX * T.hstride, 0,
                       sizeof(double) * T.hstride));
    if (T.nvecG <= 9) {  // G and the partials come from its GEMV pack
      T.gemv = true;
      continue;
    }
    T.G = s->arena.alloc<double>((size_t)T.nvecG * T.gstride);
    T.LS = s->arena.alloc<double>((size_t)T.nvecG * T.nR);
    GemmArgs a{};
    a.M = T.nR;
    a.N = T.nvecG;
    a.K = irF.ntot;
    a.A = T.P.blkF->vec + (int64_t)T.P.Flo * T.P.blkF->ld;
    a.lda = T.P.blkF->ld;
    a.a_kcontig = true;
    a.B = T.G;
    a.ldb = T.gstride;
    a.C = T.LS;
    a.ldc = T.nR;
    a.nz = 1;
    a.ksplit = 1;
    std::vector<GroupTile>& tl = pjt[T.phase];
    add_tiles(tl, (int)pj[T.phase].size(), a.M, a.N, 1, s->pj_cfg);
    pj[T.phase].push_back(a);
  }
  // ---- GEMV packs: tasks of one final block, taken phase by phase (then in
  // task order), up to 9 vectors each; a pack runs in its last task's phase
  std::vector<int> torder(s->tasks.size());
  std::iota(torder.begin(), torder.end(), 0);
  std::stable_sort(torder.begin(), torder.end(), [&](int x, int y) {
    return s->tasks[x].phase < s->tasks[y].phase;
  });
  std::map<std::pair<int, int>, int> open_unit;
  for (int ti : torder) {
    gain_sweep::Task& T = s->tasks[ti];
    if (!T.compute || !T.gemv) continue;
    const auto key = std::make_pair(T.P.jindF, T.P.gF);
    auto it = open_unit.find(key);
    if (it == open_unit.end() ||
        s->units[it->second].nvec + T.nvecG > 9) {
      gain_sweep::Unit U;
      U.blkF = T.P.blkF;
      U.ntot = T.P.bF->irrep[T.P.gF].ntot;
      U.gstride = T.gstride;
      U.Flo = T.P.Flo;
      U.Fhi = T.P.Fhi;
      // K split so that every unit's rows spread over many waves (a block
      // holds a few thousand records); chunks of at least 4096 elements
      U.ksplit = std::max(1, std::min(16, U.ntot / 4096));
      open_unit[key] = (int)s->units.size();
      s->units.push_back(U);
      it = open_unit.find(key);
    }
    gain_sweep::Unit& U = s->units[it->second];
    T.unit = it->second;
    T.vecoff = U.nvec;
    U.nvec += T.nvecG;
    U.Flo = std::min(U.Flo, T.P.Flo);
    U.Fhi = std::max(U.Fhi, T.P.Fhi);
    U.ready = std::max(U.ready, T.phase);
  }
  for (gain_sweep::Unit& U : s->units) {
    U.G = s->arena.alloc<double>((size_t)U.nvec * U.gstride);
    U.part = s->arena.alloc<double>((size_t)U.ksplit * U.nvec *
                                    (U.Fhi - U.Flo));
  }
  for (gain_sweep::Task& T : s->tasks) {
    if (T.unit < 0) continue;
    const gain_sweep::Unit& U = s->units[T.unit];
    T.G = U.G + (int64_t)T.vecoff * U.gstride;
    T.LS = U.part + T.vecoff;
    T.ready = U.ready;
  }
  for (size_t pi = 0; pi < s->phases.size(); ++pi) {
    if (pj[pi].empty()) continue;
    gain_sweep::Phase& ph = s->phases[pi];
    ph.d_pj_probs = up(s, pj[pi]);
    ph.d_pj_tiles = up(s, pjt[pi]);
    ph.n_pj_tiles = (int)pjt[pi].size();
  }
  build_launches(s);
}

// Argument tables of the downstream stages of every phase (all buffers are
// owned by the sweep, so the pointers stay valid for every replay).
void build_launches(gain_sweep* s) {
  gain_ctx* c = s->c;
  const gain_params& p = c->params;
  const int nc = c->nc;
  const int nph = (int)s->phases.size();
  std::vector<std::vector<GatherArgs>> gv(nph);
  std::vector<std::vector<ExpandTArgs>> ev(nph);
  for (gain_sweep::Task& T : s->tasks) {
    if (!T.compute) continue;
    const gain_sweep::Group& G = s->groups[T.group];
    gain_sweep::Phase& ph = s->phases[T.phase];
    Basis& bF = *T.P.bF;
    const Irrep& irF = bF.irrep[T.P.gF];
    if (T.P.bI->rotsym)
      ph.gather_rot.push_back(make_gather_rot_args(
          c, T.P, G.nvecX, 1, G.Y, T.H, s->ncol_all * nc,
          3 * s->ncol_all * nc));
    GatherArgs ga{};
    ga.nprimF = bF.nprim;
    ga.nc = nc;
    ga.jI = T.P.bI->J;
    ga.jF = bF.J;
    ga.jmax = c->jmax;
    ga.ncolI = T.P.bI->ncol;
    ga.nvec = G.nvecX;
    ga.ksplit = 1;
    ga.kvalF = bF.kval;
    ga.tauF = bF.tau;
    ga.vibF = bF.vib;
    ga.threej = c->d_threej;
    ga.Y = G.Y;
    ga.ys_alpha = s->ncol_all * nc;
    ga.ys_ks = 3 * s->ncol_all * nc;
    ga.H = T.H;
    ga.hstride = T.hstride;
    if (!T.P.bI->rotsym) gv[T.phase].push_back(ga);
    ExpandTArgs ea{};
    ea.ntot = irF.ntot;
    ea.nprim = bF.nprim;
    ea.sdeg = irF.sdeg;
    ea.tcase = irF.tcase;
    ea.case_ptr = bF.case_ptr;
    ea.case_prim = bF.case_prim;
    ea.icontr = bF.icontr;
    ea.repres = irF.repres;
    ea.H = T.H;
    ea.hstride = T.hstride;
    ea.h_of_vec = T.h_of_vec;
    ea.ideg_of_vec = T.ideg_of_vec;
    ea.nvec = T.nvecG;
    ea.G = T.G;
    ea.gstride = T.gstride;
    ev[T.phase].push_back(ea);
  }
  // every gather before any projection input: expand_T reads H of its task
  // only, so the two stages are ordered launches
  for (int pi = 0; pi < nph; ++pi) {
    gain_sweep::Phase& ph = s->phases[pi];
    add_stage(s, ph.gather, kGroupedGather, gv[pi],
              grouped_gather_table(gv[pi]));
    add_stage(s, ph.expand, kGroupedExpandT, ev[pi],
              grouped_expand_T_table(ev[pi]));
  }

  std::vector<std::vector<GemvArgs>> g4(nph), g8(nph);
  for (const gain_sweep::Unit& U : s->units) {
    GemvArgs a{};
    a.C = U.blkF->vec;
    a.ld = U.blkF->ld;
    a.ntot = U.ntot;
    a.F0 = U.Flo;
    a.nF = U.Fhi - U.Flo;
    a.G = U.G;
    a.gstride = U.gstride;
    a.nvec = U.nvec;
    a.partial = U.part;
    a.ksplit = U.ksplit;
    const bool aligned = (a.ld % 2 == 0) && (a.gstride % 2 == 0) &&
                         ((uintptr_t)a.C % 16 == 0) &&
                         ((uintptr_t)a.G % 16 == 0);
    // with many vectors per record row, more rows per wave cut the re-reads
    // of the projected vectors
    a.rows = U.nvec > 4 ? 8 : 4;
    if (!aligned)
      s->phases[U.ready].gemv_single.push_back(a);
    else if (U.nvec > 4)
      g8[U.ready].push_back(a);
    else
      g4[U.ready].push_back(a);
  }
  for (int pi = 0; pi < nph; ++pi) {
    gain_sweep::Phase& ph = s->phases[pi];
    add_stage(s, ph.gemv, kGroupedGemv4, g4[pi],
              grouped_gemv_table(g4[pi], 4));
    add_stage(s, ph.gemv, kGroupedGemv8, g8[pi],
              grouped_gemv_table(g8[pi], 8));
  }
  // algorithmic bytes of the GEMV stage: record rows inside each pack's
  // window, its projected vectors and partials (tuning getter
  // "sweep_gemv_bytes", for bandwidth figures)
  int64_t gb = 0;
  for (const gain_sweep::Unit& U : s->units)
    gb += 8 * ((int64_t)(U.Fhi - U.Flo) * U.ntot + (int64_t)U.nvec * U.ntot +
               (int64_t)U.ksplit * (U.Fhi - U.Flo) * U.nvec);
  c->last["sweep_gemv_bytes"] = gb;
  c->last["sweep_gemv_packs"] = (int64_t)s->units.size();
  c->last["sweep_phases"] = nph;

  std::vector<std::vector<LinesArgs>> lv(nph);
  for (gain_sweep::Task& T : s->tasks) {
    LinesArgs la{};
    la.nI = T.P.nI;
    la.nF = T.P.blkF->n;
    la.Flo = T.compute ? T.P.Flo : 0;
    la.Fhi = T.compute ? T.P.Fhi : 0;
    if (T.unit >= 0) {  // rows of the pack's window
      la.Flo = s->units[T.unit].Flo;
      la.Fhi = s->units[T.unit].Fhi;
    }
    la.maxdeg = c->maxdeg;
    la.nvec_per_i = T.nvpi;
    la.slot_of_vec = T.slot_of_vec;
    la.src = T.LS;
    if (T.unit >= 0) {  // pack partials [ks][F][vec], this task's columns
      const gain_sweep::Unit& U = s->units[T.unit];
      la.ksplit = U.ksplit;
      la.src_ks = (int64_t)(U.Fhi - U.Flo) * U.nvec;
      la.src_f = U.nvec;
      la.src_v = 1;
    } else {  // [vec][F]
      la.ksplit = 1;
      la.src_f = 1;
      la.src_v = T.nR;
    }
    la.eI = T.eI;
    la.levelI = T.levelI;
    la.okI = T.okI;
    la.eF = T.P.blkF->d_energy;
    la.same_block = (T.P.jindI == T.P.jindF && T.P.gI == T.P.gF);
    la.block_ok = T.P.block_ok;
    la.jI = T.P.bI->J;
    la.jF = T.P.bF->J;
    la.ndegI = p.sym_degen[T.P.gI];
    la.ndegF = p.sym_degen[T.P.gF];
    la.reduced = p.reduced;
    la.gnsF = p.gns[T.P.gF];
    la.zpe = p.zpe;
    la.threshold = p.threshold;
    la.beta = kPlanck * kVellgt / (kBoltz * p.temperature);
    la.partition = p.partition;
    la.fw0 = p.freq_window[0];
    la.fw1 = p.freq_window[1];
    la.eu0 = p.erange_upper[0];
    la.eu1 = p.erange_upper[1];
    la.isym_do_F = p.isym_do[T.P.gF];
    la.ls = T.out.ls;
    la.S = T.out.S;
    la.A = T.out.A;
    la.I = T.out.I;
    la.flags = T.out.flags;
    lv[T.ready].push_back(la);
  }
  for (int pi = 0; pi < nph; ++pi)
    add_stage(s, s->phases[pi].lines, kGroupedLines, lv[pi],
              grouped_lines_table(lv[pi]));
}

// Everything downstream of one phase's mu X, on stream st: gathers,
// transposed expansions, the phase's grouped projection GEMM, the GEMV packs
// that became complete, the epilogue of their tasks.
void downstream(gain_sweep* s, const gain_sweep::Phase& ph, hipStream_t st) {
  run_stage(ph.gather, st);
  for (const GatherRotArgs& a : ph.gather_rot) launch_gather_rotsym(a, st);
  run_stage(ph.expand, st);
  if (ph.n_pj_tiles > 0)
    gemm_f64_grouped(ph.d_pj_probs, ph.d_pj_tiles, ph.n_pj_tiles, true,
                     s->pj_cfg, s->grid, st);
  run_stage(ph.gemv, st);
  for (const GemvArgs& a : ph.gemv_single) launch_gemv(a, st);
  run_stage(ph.lines, st);
}

// The launch sequence of one sweep (kernels and event record/wait pairs):
// per phase, its scatter and mu X on st, then its downstream stages on the
// side stream after an event; finally st waits for the side stream.  With
// one phase everything runs on st.
void enqueue(gain_sweep* s, hipStream_t st) {
  const bool two = s->phases.size() > 1;
  hipStream_t ds = two ? s->side : st;
  for (const gain_sweep::Phase& ph : s->phases) {
    run_stage(ph.scatter, st);
    products(s, ph, st);
    if (two) {
      GAIN_HIP(hipEventRecord(ph.done, st));
      GAIN_HIP(hipStreamWaitEvent(ds, ph.done, 0));
    }
    downstream(s, ph, ds);
  }
  if (two) {
    GAIN_HIP(hipEventRecord(s->join, ds));
    GAIN_HIP(hipStreamWaitEvent(st, s->join, 0));
  }
}

void release(gain_sweep* s) {
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  for (gain_sweep::Phase& ph : s->phases)
    if (ph.done) (void)hipEventDestroy(ph.done);
  if (s->join) (void)hipEventDestroy(s->join);
  if (s->side) (void)hipStreamDestroy(s->side);
  delete s;
}

}  // namespace

extern "C" int gain_sweep_create(gain_ctx* c, const gain_task* tasks,
                                 int ntasks, int precision,
                                 gain_sweep** out) {
  gain_sweep* s = nullptr;
  int rc = guarded([&] {
    need(c);
    GAIN_CHECK_ARG(out && tasks && ntasks > 0, "bad sweep arguments");
    *out = nullptr;
    GAIN_CHECK_ARG(precision == GAIN_PREC_FP64 ||
                       precision == GAIN_PREC_FP32 ||
                       slices_of(precision) > 0 ||
                       i8_slices_of(precision) > 0 ||
                       crt_moduli_of(precision) > 0,
                   "sweeps run FP64, FP32, split-BF16, split-INT8 or "
                   "CRT-form products (precision %d)",
                   precision);
    if (!c->params_set) fail(GAIN_ERR_STATE, "gain_set_params not called");
    need_mu(c);
    DeviceGuard g(c->device);
    s = new gain_sweep();
    s->c = c;
    s->slices = i8_slices_of(precision);
    s->crt = crt_moduli_of(precision);
    s->f32 = precision == GAIN_PREC_FP32;
    s->f32_cfg = c->f32_cfg;
    s->bf16 = slices_of(precision);
    if (c->force_cfg[0] >= 0) s->pj_cfg = c->force_cfg[0];
    GAIN_HIP(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    GAIN_HIP(hipEventCreateWithFlags(&s->join, hipEventDisableTiming));
    build(s, tasks, ntasks);
    GAIN_HIP(hipStreamSynchronize(c->stream));
    enqueue(s, c->stream);  // eager run: validates every launch
    GAIN_HIP(hipStreamSynchronize(c->stream));
    GAIN_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue(s, c->stream);
    } catch (...) {
      hipGraph_t dead = nullptr;
      (void)hipStreamEndCapture(c->stream, &dead);
      if (dead) (void)hipGraphDestroy(dead);
      throw;
    }
    GAIN_HIP(hipStreamEndCapture(c->stream, &s->graph));
    GAIN_HIP(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
    s->gen = c->gen;
    *out = s;
  });
  if (rc != GAIN_OK && s) {
    (void)hipStreamSynchronize(c->stream);
    if (s->side) (void)hipStreamSynchronize(s->side);
    release(s);
  }
  return rc;
}

extern "C" int gain_sweep_run(gain_sweep* s) {
  return guarded([&] {
    GAIN_CHECK_ARG(s && s->exec, "null sweep");
    if (s->gen != s->c->gen)
      fail(GAIN_ERR_STATE,
           "sweep invalidated: the dipole, a state block or a basis table "
           "it captured was replaced or changed after gain_sweep_create; "
           "create the sweep again");
    DeviceGuard g(s->c->device);
    if (s->c->sweep_eager)
      enqueue(s, s->c->stream);  // the same sequence, launched one by one
    else
      GAIN_HIP(hipGraphLaunch(s->exec, s->c->stream));
  });
}

extern "C" int gain_sweep_destroy(gain_sweep* s) {
  return guarded([&] {
    if (!s) return;
    DeviceGuard g(s->c->device);
    (void)hipStreamSynchronize(s->c->stream);
    if (s->side) (void)hipStreamSynchronize(s->side);
    release(s);
  });
}
