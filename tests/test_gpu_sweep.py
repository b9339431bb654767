"""Mixed-J sweeps: uneven (J, irrep) batches against every final block with
dJ in {-1, 0, +1}, captured once as a HIP graph and replayed.  Every replay
must reproduce the unbatched plain FP64 path: transition lists identical,
S within 1e-12 on strong lines.  Between replays the context does other
(eager) work, which must not disturb the sweep."""
import numpy as np
import pytest

from gain_engine import synth
from gain_engine.engine import Engine, Sweep

pytestmark = pytest.mark.gpu


def _tasks(prob, eng, rng, max_per_block):
    import torch
    tasks, refs = [], []
    pairs = {g: eng.gamma_pair(g) for g in range(prob.params.nrepres)}
    for (jI, gI), blk in sorted(prob.blocks.items()):
        n = int(rng.integers(1, max_per_block + 1))
        levels = sorted(rng.choice(blk.n, size=min(n, blk.n),
                                   replace=False).tolist())
        for (jF, gF), bF in sorted(prob.blocks.items()):
            if gF != pairs[gI] or abs(prob.bases[jF].J - prob.bases[jI].J) > 1:
                continue
            out = {k: torch.zeros((len(levels), bF.n), dtype=torch.float64,
                                  device="cuda:0") for k in ("S", "A", "I")}
            out["flags"] = torch.zeros((len(levels), bF.n),
                                       dtype=torch.uint8, device="cuda:0")
            tasks.append((jI, gI, levels, jF, gF,
                          {k: v.data_ptr() for k, v in out.items()}))
            refs.append(out)
    return tasks, refs


def _check(eng, tasks, outs, rel=1e-12):
    n = 0
    for (jI, gI, levels, jF, gF, _), out in zip(tasks, outs):
        flags = out["flags"].cpu().numpy()
        S = out["S"].cpu().numpy()
        for i, lv in enumerate(levels):
            r = eng.plain_lines(jI, gI, lv, jF, gF, want_ls=False)
            assert np.array_equal(r["flags"], flags[i]), (jI, gI, jF, gF, lv)
            ref = r["S"]
            if ref.max() <= 0:
                continue
            strong = ref > 1e-6 * ref.max()
            d = np.abs(S[i] - ref)
            assert (d[strong] <= rel * ref[strong]).all()
            assert (d <= rel * ref.max()).all()
            n += int((flags[i] & 1).sum())
    return n


@pytest.mark.parametrize("cfgs,phases", [((-1, -1), 2), ((0, 7), 1),
                                         ((9, 3), 3), ((-1, -1), 16)])
def test_mixed_j_sweep_graph_replays_match_unbatched(cfgs, phases):
    """Also over the pipeline phases: 1 (one stream), 2 (the default), 3
    and 16 (one batch per phase: each phase's projections on the second
    stream beside the next phase's mu X)."""
    import torch
    prob = synth.make_problem([5, 6, 7, 8], 12, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rng = np.random.default_rng(11)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("proj_cfg", cfgs[0])
        eng.set_tuning("half_ls_cfg", cfgs[1])
        eng.set_tuning("sweep_phases", phases)
        tasks, outs = _tasks(prob, eng, rng, max_per_block=24)
        assert sum(len(t[2]) for t in tasks) > 256
        sw = Sweep(eng, tasks)
        total = None
        for rep in range(3):
            for o in outs:
                for v in o.values():
                    v.zero_()
            torch.cuda.synchronize()
            sw.run()
            eng.synchronize()
            n = _check(eng, tasks, outs)
            assert n > 0 and (total is None or n == total)
            total = n
            # unrelated eager work on the same context between replays
            blk = next(iter(prob.blocks))
            eng.batched_lines(blk[0], blk[1], [0, 1], blk[0] + 1
                              if blk[0] + 1 < len(prob.bases) else blk[0],
                              eng.gamma_pair(blk[1]))
        sw.close()


@pytest.mark.parametrize("precision,rel,phases", [(8, 1e-7, 4),
                                                   (8, 1e-7, 1),
                                                   (9, 1e-9, 16)])
def test_mixed_j_sweep_crt(precision, rel, phases):
    """A sweep whose mu X products run in the CRT form (12 or 14 moduli, the
    sweep's own residues of mu, X cut into residue planes per phase inside
    the graph) against the unbatched plain FP64 path: transition lists
    identical, S within `rel` on strong lines, on two replays with eager
    CRT and FP64 work on the context between them."""
    import torch
    prob = synth.make_problem([5, 6, 7, 8], 12, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rng = np.random.default_rng(7)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("sweep_phases", phases)
        tasks, outs = _tasks(prob, eng, rng, max_per_block=24)
        sw = Sweep(eng, tasks, precision=precision)
        for rep in range(2):
            for o in outs:
                for v in o.values():
                    v.zero_()
            torch.cuda.synchronize()
            sw.run()
            eng.synchronize()
            assert _check(eng, tasks, outs, rel=rel) > 0
            blk = next(iter(prob.blocks))
            jF = blk[0] + 1 if blk[0] + 1 < len(prob.bases) else blk[0]
            eng.batched_lines(blk[0], blk[1], [0, 1], jF,
                              eng.gamma_pair(blk[1]), precision=precision)
            eng.batched_lines(blk[0], blk[1], [0, 1], jF,
                              eng.gamma_pair(blk[1]))
        sw.close()


@pytest.mark.parametrize("precision", [5, 6])
def test_mixed_j_sweep_split_int8(precision):
    """A sweep whose mu X products run as split-INT8 (6 or 7 slices, the
    sweep's own int8 copy of mu, X sliced inside the graph) against the
    unbatched plain FP64 path: transition lists identical, S within 1e-7
    (6 slices) or 1e-9 (7 slices) on strong lines, on two replays with
    eager work between them."""
    import torch
    prob = synth.make_problem([5, 6, 7, 8], 12, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rng = np.random.default_rng(5)
    with Engine.from_problem(prob) as eng:
        tasks, outs = _tasks(prob, eng, rng, max_per_block=24)
        sw = Sweep(eng, tasks, precision=precision)
        for rep in range(2):
            for o in outs:
                for v in o.values():
                    v.zero_()
            torch.cuda.synchronize()
            sw.run()
            eng.synchronize()
            assert _check(eng, tasks, outs,
                          rel=1e-7 if precision == 5 else 1e-9) > 0
            # eager split-INT8 and FP64 work on the context in between
            blk = next(iter(prob.blocks))
            jF = blk[0] + 1 if blk[0] + 1 < len(prob.bases) else blk[0]
            eng.batched_lines(blk[0], blk[1], [0, 1], jF,
                              eng.gamma_pair(blk[1]), precision=4)
            eng.batched_lines(blk[0], blk[1], [0, 1], jF,
                              eng.gamma_pair(blk[1]))
        sw.close()


@pytest.fixture(scope="module")
def mixed_full():
    """The full mixed-J shape: 256 initial states drawn from J_I in [5, 30]
    (Nc = 3000, uneven blocks N_J = (2J+1) Nc / 4, 2000 states per block,
    48 GB of records in HBM), every dJ in {-1, 0, +1} final block, with the
    unbatched plain FP64 lines of every task as the reference."""
    import torch
    from gain_engine import workload
    w = workload.build_mixed()
    try:
        eng = w.eng
        tasks = w.tasks()
        refs = []
        for (jI, gI, levels, jF, gF) in tasks:
            rs = [eng.plain_lines(jI, gI, lv, jF, gF, want_ls=False)
                  for lv in levels]
            refs.append((np.stack([r["flags"] for r in rs]),
                         np.stack([r["S"] for r in rs])))
        yield w, tasks, refs
    finally:
        torch.cuda.synchronize()
        w.close()


def _full_sweep(w, tasks, refs, precision, check):
    import torch
    eng = w.eng
    outs, full = [], []
    for (jI, gI, levels, jF, gF) in tasks:
        nF = w.prob.blocks[(jF, gF)].n
        o = {"S": torch.zeros((len(levels), nF), dtype=torch.float64,
                              device="cuda:0"),
             "flags": torch.zeros((len(levels), nF), dtype=torch.uint8,
                                  device="cuda:0")}
        outs.append(o)
        full.append((jI, gI, levels, jF, gF,
                     {k: v.data_ptr() for k, v in o.items()}))
    torch.cuda.synchronize()
    sw = Sweep(eng, full, precision=precision)
    try:
        for rep in range(2):
            for o in outs:
                for v in o.values():
                    v.zero_()
            torch.cuda.synchronize()
            sw.run()
            eng.synchronize()
            nlines = 0
            for t, o, (rfl, rS) in zip(tasks, outs, refs):
                fl = o["flags"].cpu().numpy()
                S = o["S"].cpu().numpy()
                assert np.array_equal(fl, rfl), t
                for i in range(len(t[2])):
                    if rS[i].max() > 0:
                        check(S[i], rS[i], t)
                nlines += int((fl & 1).sum())
            assert nlines > 100000
    finally:
        sw.close()


@pytest.mark.slow
@pytest.mark.parametrize("phases", [2, 1])
def test_mixed_j_sweep_full_shape_matches_unbatched(mixed_full, phases):
    """One captured FP64 sweep of the full shape (pipelined over 2 phases,
    and on one stream) replayed twice, against the unbatched plain FP64
    path: transition flags identical; S within 1e-12 relative (the
    config-2 bar) on every line with S > 1e-4 S_max of its initial state
    (measured: 3.4e-13), and within the condition-normalised FP64 bound
    |dS| <= 2e-12 S + 2e-14 sqrt(S S_max) on the weaker lines, whose
    relative error grows with the cancellation in their sums (the two paths
    add in different orders; 3.3e-12 above 1e-6 S_max)."""
    w, tasks, refs = mixed_full
    w.eng.set_tuning("sweep_phases", phases)

    def check(S, ref, t):
        d = np.abs(S - ref)
        assert (d <= 2e-12 * ref + 2e-14 * np.sqrt(ref * ref.max())).all(), t
        strong = ref > 1e-4 * ref.max()
        assert (d[strong] <= 1e-12 * ref[strong]).all(), t
    _full_sweep(w, tasks, refs, 0, check)


@pytest.mark.slow
def test_mixed_j_sweep_full_shape_crt12(mixed_full):
    """The same full-shape sweep with mu X in the CRT form (12 moduli),
    pipelined over 4 phases: transition flags identical to the plain FP64
    path, S within 1e-7 on strong lines (S > 1e-6 S_max of its initial
    state) and within 1e-7 S_max on every line."""
    w, tasks, refs = mixed_full
    w.eng.set_tuning("sweep_phases", 4)

    def check(S, ref, t):
        d = np.abs(S - ref)
        strong = ref > 1e-6 * ref.max()
        assert (d[strong] <= 1e-7 * ref[strong]).all(), t
        assert (d <= 1e-7 * ref.max()).all(), t
    _full_sweep(w, tasks, refs, 8, check)


@pytest.mark.parametrize("change", ["dipole", "block", "combination"])
def test_sweep_refuses_replay_after_captured_data_changes(change):
    """A sweep's graph holds raw device pointers: once the dipole is
    uploaded again, a block it uses is replaced, or mu is changed in place,
    gain_sweep_run returns GAIN_ERR_STATE without replaying."""
    import torch
    from gain_engine import _lib
    prob = synth.make_problem([2, 3], 12, "c2v4")
    rng = np.random.default_rng(3)
    with Engine.from_problem(prob) as eng:
        tasks, outs = _tasks(prob, eng, rng, 3)
        sw = Sweep(eng, tasks)
        sw.run()
        eng.synchronize()
        for o in outs:
            o["flags"].fill_(7)
        torch.cuda.synchronize()
        if change == "dipole":
            eng.upload_dipole(prob.mu)
        elif change == "block":
            (j, g), blk = sorted(prob.blocks.items())[0]
            eng.upload_block(j, g, blk.energy, blk.vectors)
        else:
            M = torch.as_tensor(prob.mu.reshape(1, -1), device="cuda:0")
            eng.set_dipole_combination(np.ones(1), M.data_ptr(), prob.nc)
        with pytest.raises(_lib.GainError) as e:
            sw.run()
        assert e.value.code == _lib.ERR_STATE
        eng.synchronize()
        # nothing was replayed: the outputs keep the marker
        assert all(bool((o["flags"] == 7).all()) for o in outs)
        sw.close()
        # a sweep created afterwards runs
        sw2 = Sweep(eng, tasks)
        sw2.run()
        eng.synchronize()
        assert _check(eng, tasks, outs) > 0
        sw2.close()


@pytest.mark.parametrize("reduced", [True, False])
@pytest.mark.parametrize("precision", [0, 8])
def test_mixed_j_sweep_td(reduced, precision):
    """Td symmetry (3-dimensional irreps; unreduced: nine projected vectors
    per F1 <-> F2 initial state, the largest GEMV pack) through a captured
    sweep, against the unbatched plain FP64 path."""
    prob = synth.make_problem([2, 3, 4], 20, "td5")
    prob.params.reduced = reduced
    prob.params.erange_lower = (0.0, 1e9)
    rng = np.random.default_rng(5)
    with Engine.from_problem(prob) as eng:
        tasks, outs = _tasks(prob, eng, rng, max_per_block=6)
        sw = Sweep(eng, tasks, precision=precision)
        for _ in range(2):
            sw.run()
            eng.synchronize()
            assert _check(eng, tasks, outs,
                          rel=1e-12 if precision == 0 else 1e-7) > 0
        sw.close()


@pytest.mark.parametrize("precision,rel", [(1, 1e-3), (2, 1e-2), (3, 1e-4)])
def test_mixed_j_sweep_fp32_and_split_bf16(precision, rel):
    """A sweep whose mu X products run on the FP32 matrix cores (FP32
    products, FP64 accumulation) or as split-BF16 with 3 or 4 slices (the
    sweep's own bf16 slices of mu, X sliced inside the graph) against the
    unbatched plain FP64 path: transition lists identical, S within `rel`
    on strong lines, on two replays with eager work between them."""
    import torch
    prob = synth.make_problem([5, 6, 7, 8], 12, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    rng = np.random.default_rng(11)
    with Engine.from_problem(prob) as eng:
        eng.set_tuning("sweep_phases", 2)
        tasks, outs = _tasks(prob, eng, rng, max_per_block=24)
        sw = Sweep(eng, tasks, precision=precision)
        for rep in range(2):
            for o in outs:
                for v in o.values():
                    v.zero_()
            torch.cuda.synchronize()
            sw.run()
            eng.synchronize()
            assert _check(eng, tasks, outs, rel=rel) > 0
            blk = next(iter(prob.blocks))
            jF = blk[0] + 1 if blk[0] + 1 < len(prob.bases) else blk[0]
            eng.batched_lines(blk[0], blk[1], [0, 1], jF,
                              eng.gamma_pair(blk[1]), precision=precision)
        sw.close()
