// Mixed-J sweeps: many (initial batch -> final block) tasks in one launch
// sequence captured as a HIP graph.
//
// Tasks with the same initial batch (J index, irrep, levels) share X and
// Y = mu X.  The X of every batch sits in one column range of a single
// matrix, so mu X for the whole sweep is one GEMM (tile shape from the
// autotuner; FP64, or split-INT8 on the sweep's own int8 copy of mu); the
// projections C_F G of the tasks with many vectors run as one persistent
// grouped launch over the tiles of all of them.
// Tasks that project at most 9 vectors go through the FP64 GEMV instead (a
// 128-wide GEMM tile would be mostly padding); such tasks that target the
// same final block are packed into one GEMV of up to 9 vectors over the
// union of their energy windows, so the block's records are read once per
// pack rather than once per task, and every pack of the sweep runs in one
// grouped GEMV launch.  The per-batch and per-task kernels (scatter, gather,
// transposed expansion, epilogue) also run as one grouped launch per stage
// over argument tables built at creation.
// Everything a replay touches is owned by the sweep (sized at creation), so
// replays stay valid whatever the context does in between.  The launch
// sequence contains kernels only (no memset or memcpy nodes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <tuple>
#include <vector>

#include "engine_internal.hpp"

using namespace gain;

struct gain_sweep {
  gain_ctx* c = nullptr;
  DeviceArena arena;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  struct Group {
    Plan P;
    int nvecX = 0, ncolTot = 0;
    double* X = nullptr;
    double* Y = nullptr;
    int64_t* c_off = nullptr;
    int* ideg = nullptr;
  };
  struct Task {
    Plan P;
    int group = 0;
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
    double* G = nullptr;
    double* part = nullptr;  // [ksplit][Fhi-Flo][nvec]
  };
  std::vector<Group> groups;
  std::vector<Task> tasks;
  std::vector<Unit> units;
  // one grouped launch per stage: argument table + workgroup offsets
  struct Launch {
    int kind = 0;
    const void* args = nullptr;
    const int* start = nullptr;
    int n = 0, blocks = 0;
  };
  std::vector<Launch> stage_scatter, stage_gather, stage_expand, stage_gemv,
      stage_lines;
  std::vector<GatherRotArgs> gather_rot;  // symmetric-top tasks, one launch
                                          // each
  std::vector<GemvArgs> gemv_single;  // units whose layout the grouped
                                      // GEMV cannot take
  // mu X of every batch: one GEMM over [X of batch 0 | X of batch 1 | ...]
  GemmArgs hl{};
  bool has_hl = false;
  int64_t ncol_all = 0;
  // split-INT8 (GAIN_PREC_I8X*) or CRT-form (GAIN_PREC_I8CRT*) mu X: the
  // sweep's own int8 copy of mu, X cut into planes at every replay
  int slices = 0;
  int crt = 0;  // moduli of the CRT form, 0 otherwise
  I8GemmArgs hl8{};
  int64_t x8_plane = 0;
  GemmArgs* d_pj_probs = nullptr;
  GroupTile* d_pj_tiles = nullptr;
  int n_pj_tiles = 0, pj_cfg = 5;
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
  // ---- groups: X, Y (column ranges of one matrix), scatter arrays
  std::vector<int64_t> col0(s->groups.size(), 0);
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
    col0[g] = s->ncol_all;
    s->ncol_all += G.ncolTot;
  }
  if (s->ncol_all > 0) {
    GAIN_CHECK_ARG(s->ncol_all < (1ll << 31), "sweep: too many columns");
    double* X = s->arena.alloc<double>((size_t)s->ncol_all * nc);
    double* Y = s->arena.alloc<double>((size_t)3 * s->ncol_all * nc);
    for (size_t g = 0; g < s->groups.size(); ++g) {
      gain_sweep::Group& G = s->groups[g];
      if (G.nvecX == 0) continue;
      G.X = X + col0[g] * nc;
      G.Y = Y + col0[g] * nc;  // alpha planes are ncol_all * nc apart
    }
    GemmArgs& a = s->hl;
    a.M = nc;
    a.N = (int)s->ncol_all;
    a.K = nc;
    a.A = c->d_mu;
    a.lda = nc;
    a.sAz = (int64_t)nc * nc;
    a.a_kcontig = false;
    a.B = X;
    a.ldb = nc;
    a.C = Y;
    a.ldc = nc;
    a.sCz = s->ncol_all * nc;
    a.nz = 3;
    a.ksplit = 1;
    s->has_hl = true;
    if (s->crt > 0) {
      const int nm = s->crt;
      double* muT = s->arena.alloc<double>((size_t)3 * nc * nc);
      transpose_mu(c->d_mu, nc, muT, c->stream);
      const int64_t Kp = i8_kpad(nc), Rm = i8_rows_pad(nc);
      const int beta = crt_beta(nm, Kp);
      int8_t* m8 = s->arena.alloc<int8_t>((size_t)nm * 3 * Rm * Kp);
      int* me = s->arena.alloc<int>((size_t)3 * Rm);
      for (int al = 0; al < 3; ++al)
        slice_rows_crt(nm, beta, muT + (size_t)al * nc * nc, nc, nc, nc,
                       m8 + al * Rm * Kp, 3 * Rm * Kp, me + al * Rm,
                       c->stream);
      GAIN_HIP(hipStreamSynchronize(c->stream));
      s->arena.release(muT);
      const int64_t Rx = i8_rows_pad(s->ncol_all);
      s->x8_plane = Rx * Kp;
      I8GemmArgs& g = s->hl8;
      g.S = nm;
      g.beta = beta;
      g.M = nc;
      g.N = (int)s->ncol_all;
      g.row0 = 0;
      g.Kp = Kp;
      g.A = m8;
      g.a_plane = 3 * Rm * Kp;
      g.a_z = Rm * Kp;
      g.eA = me;
      g.ea_z = Rm;
      g.B = s->arena.alloc<int8_t>((size_t)nm * s->x8_plane);
      g.b_plane = s->x8_plane;
      g.eB = s->arena.alloc<int>((size_t)Rx);
      g.C = Y;
      g.ldc = nc;
      g.c_z = s->ncol_all * nc;
      g.nz = 3;
      int64_t rb = 0;  // residue buffer for every tile configuration
      for (int cfg = 0; cfg < kCrtConfigs; ++cfg) {
        g.cfg = cfg;
        rb = std::max(rb, gemm_crt_res_bytes(g));
      }
      g.R = s->arena.alloc<int8_t>((size_t)rb);
      // cut X once so the tuner times real operands
      slice_rows_crt(nm, beta, X, nc, (int)s->ncol_all, nc, (int8_t*)g.B,
                     s->x8_plane, (int*)g.eB, c->stream);
      g.cfg = crt_cfg_of(c, 1, g);
    } else if (s->slices == 0) {
      a.cfg = gemm_cfg(c, 1, a);  // tuned here: replays must not time
    } else {
      const int S = s->slices;
      double* muT = s->arena.alloc<double>((size_t)3 * nc * nc);
      transpose_mu(c->d_mu, nc, muT, c->stream);
      const int64_t Kp = i8_kpad(nc), Rm = i8_rows_pad(nc);
      int8_t* m8 = s->arena.alloc<int8_t>((size_t)S * 3 * Rm * Kp);
      int* me = s->arena.alloc<int>((size_t)3 * Rm);
      for (int al = 0; al < 3; ++al)
        slice_rows_i8(S, muT + (size_t)al * nc * nc, nc, nc, nc,
                      m8 + al * Rm * Kp, 3 * Rm * Kp, me + al * Rm,
                      c->stream);
      GAIN_HIP(hipStreamSynchronize(c->stream));
      s->arena.release(muT);
      const int64_t Rx = i8_rows_pad(s->ncol_all);
      s->x8_plane = Rx * Kp;
      I8GemmArgs& g = s->hl8;
      g.S = S;
      g.M = nc;
      g.N = (int)s->ncol_all;
      g.row0 = 0;
      g.Kp = Kp;
      g.A = m8;
      g.a_plane = 3 * Rm * Kp;
      g.a_z = Rm * Kp;
      g.eA = me;
      g.ea_z = Rm;
      g.B = s->arena.alloc<int8_t>((size_t)S * s->x8_plane);
      g.b_plane = s->x8_plane;
      g.eB = s->arena.alloc<int>((size_t)Rx);
      g.C = Y;
      g.ldc = nc;
      g.c_z = s->ncol_all * nc;
      g.c_ks = 3 * s->ncol_all * nc;
      g.nz = 3;
      int ks = 1;
      // slice once so the tuner times real operands
      slice_rows_i8(S, X, nc, (int)s->ncol_all, nc, (int8_t*)g.B,
                    s->x8_plane, (int*)g.eB, c->stream);
      g.cfg = i8_cfg_split(c, 1, g, 1, &ks);
      g.ksplit = ks;
    }
  }
  // ---- tasks: H, G, LS, per-pair arrays, projection problems
  std::vector<GemmArgs> pj;
  std::vector<GroupTile> pjt;
  for (gain_sweep::Task& T : s->tasks) {
    const gain_sweep::Group& G = s->groups[T.group];
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
    add_tiles(pjt, (int)pj.size(), a.M, a.N, 1, s->pj_cfg);
    pj.push_back(a);
  }
  // ---- GEMV packs: tasks of one final block, in task order, up to 9
  // vectors each
  std::map<std::pair<int, int>, int> open_unit;
  for (gain_sweep::Task& T : s->tasks) {
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
  }
  if (!pj.empty()) {
    s->d_pj_probs = up(s, pj);
    s->d_pj_tiles = up(s, pjt);
    s->n_pj_tiles = (int)pjt.size();
  }
  build_launches(s);
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

// Argument tables of every stage of the launch sequence (all buffers are
// owned by the sweep, so the pointers stay valid for every replay).
void build_launches(gain_sweep* s) {
  gain_ctx* c = s->c;
  const gain_params& p = c->params;
  const int nc = c->nc;
  std::vector<ScatterArgs> sc;
  for (gain_sweep::Group& G : s->groups) {
    if (G.nvecX == 0) continue;
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
  add_stage(s, s->stage_scatter, kGroupedScatter, sc,
            grouped_scatter_table(sc));

  std::vector<GatherArgs> gv;
  std::vector<ExpandTArgs> ev;
  for (gain_sweep::Task& T : s->tasks) {
    if (!T.compute) continue;
    const gain_sweep::Group& G = s->groups[T.group];
    Basis& bF = *T.P.bF;
    const Irrep& irF = bF.irrep[T.P.gF];
    if (T.P.bI->rotsym)
      s->gather_rot.push_back(make_gather_rot_args(
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
    if (!T.P.bI->rotsym) gv.push_back(ga);
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
    ev.push_back(ea);
  }
  // every gather before any projection input: expand_T reads H of its task
  // only, so the two stages are ordered launches
  add_stage(s, s->stage_gather, kGroupedGather, gv, grouped_gather_table(gv));
  add_stage(s, s->stage_expand, kGroupedExpandT, ev,
            grouped_expand_T_table(ev));

  std::vector<GemvArgs> g4, g8;
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
      s->gemv_single.push_back(a);
    else if (U.nvec > 4)
      g8.push_back(a);
    else
      g4.push_back(a);
  }
  add_stage(s, s->stage_gemv, kGroupedGemv4, g4, grouped_gemv_table(g4, 4));
  add_stage(s, s->stage_gemv, kGroupedGemv8, g8, grouped_gemv_table(g8, 8));
  // algorithmic bytes of the GEMV stage: record rows inside each pack's
  // window, its projected vectors and partials (tuning getter
  // "sweep_gemv_bytes", for bandwidth figures)
  int64_t gb = 0;
  for (const gain_sweep::Unit& U : s->units)
    gb += 8 * ((int64_t)(U.Fhi - U.Flo) * U.ntot + (int64_t)U.nvec * U.ntot +
               (int64_t)U.ksplit * (U.Fhi - U.Flo) * U.nvec);
  c->last["sweep_gemv_bytes"] = gb;
  c->last["sweep_gemv_packs"] = (int64_t)s->units.size();

  std::vector<LinesArgs> lv;
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
    lv.push_back(la);
  }
  add_stage(s, s->stage_lines, kGroupedLines, lv, grouped_lines_table(lv));
}

void run_stage(const std::vector<gain_sweep::Launch>& v, hipStream_t st) {
  for (const gain_sweep::Launch& L : v)
    launch_grouped(L.kind, L.args, L.start, L.n, L.blocks, st);
}

// The launch sequence of one sweep (kernels only): scatter of every batch,
// mu X of every batch as one GEMM, gathers, transposed expansions, the
// grouped projection GEMM and GEMVs, the epilogue of every task.
void enqueue(gain_sweep* s, hipStream_t st) {
  run_stage(s->stage_scatter, st);
  if (s->has_hl && s->slices == 0 && s->crt == 0) gemm_f64(s->hl, st);
  if (s->has_hl && s->crt > 0) {
    slice_rows_crt(s->crt, s->hl8.beta, s->hl.B, s->c->nc, (int)s->ncol_all,
                   s->c->nc, (int8_t*)s->hl8.B, s->x8_plane,
                   (int*)s->hl8.eB, st);
    gemm_crt_i8(s->hl8, st);
  }
  if (s->has_hl && s->slices > 0) {
    slice_rows_i8(s->slices, s->hl.B, s->c->nc, (int)s->ncol_all, s->c->nc,
                  (int8_t*)s->hl8.B, s->x8_plane, (int*)s->hl8.eB, st);
    gemm_sliced_i8(s->hl8, st);
  }
  run_stage(s->stage_gather, st);
  for (const GatherRotArgs& a : s->gather_rot) launch_gather_rotsym(a, st);
  run_stage(s->stage_expand, st);
  gemm_f64_grouped(s->d_pj_probs, s->d_pj_tiles, s->n_pj_tiles, true,
                   s->pj_cfg, s->grid, st);
  run_stage(s->stage_gemv, st);
  for (const GemvArgs& a : s->gemv_single) launch_gemv(a, st);
  run_stage(s->stage_lines, st);
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
                       i8_slices_of(precision) > 0 ||
                       crt_moduli_of(precision) > 0,
                   "sweeps run FP64, split-INT8 or CRT-form products "
                   "(precision %d)",
                   precision);
    if (!c->params_set) fail(GAIN_ERR_STATE, "gain_set_params not called");
    if (!c->d_mu) fail(GAIN_ERR_STATE, "dipole not uploaded");
    DeviceGuard g(c->device);
    s = new gain_sweep();
    s->c = c;
    s->slices = i8_slices_of(precision);
    s->crt = crt_moduli_of(precision);
    if (c->force_cfg[0] >= 0) s->pj_cfg = c->force_cfg[0];
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
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    delete s;
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
    GAIN_HIP(hipGraphLaunch(s->exec, s->c->stream));
  });
}

extern "C" int gain_sweep_destroy(gain_sweep* s) {
  return guarded([&] {
    if (!s) return;
    DeviceGuard g(s->c->device);
    (void)hipStreamSynchronize(s->c->stream);
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    delete s;
  });
}
