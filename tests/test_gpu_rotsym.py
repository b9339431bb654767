"""Symmetric-top (ROTSYM) bases through the engine: the half line strength
of dipole_GPU.cu:641-723 (GpuManager::ExecuteRotSymHalfLs,
GpuManager.cpp:547-600, chosen by rotsym_do at :612-618) as the mu X
products of every precision followed by the contraction with the
rotational table, then the same projection and epilogue as K-block bases.
Against the oracle's ROTSYM loop (pinned in test_rotsym.py against the
K-block loop carried into the symmetrised basis): transition lists and
emitted sets bit-exact, S, A, I within 1e-12 (FP64) or S within 1e-7 on
strong lines (CRT, 12 moduli)."""
import ctypes as C

import numpy as np
import pytest

from gain_engine import _lib, synth
from gain_engine.engine import Engine, Sweep
from oracle import oracle as orc

from helpers import compare_lines, engine_lines, oracle_map

pytestmark = pytest.mark.gpu

CASES = [("C3V(M)", "c1", [2, 3], 6), ("C3V(M)", "c3v3", [3, 4], 6),
         ("TD(M)", "td5", [2, 3], 10), ("D2H(M)", "c2v4", [1, 2], 8),
         ("D3H(M)", "c1", [4, 5], 3)]


def _problem(group, scheme, jvals, nc):
    prob = synth.make_problem(jvals, nc, scheme, rotsym=group)
    prob.params.erange_lower = (0.0, 1e9)
    return prob


@pytest.mark.parametrize("mode,precision", [("plain", 0), ("batched", 0),
                                            ("batched", 8)])
@pytest.mark.parametrize("group,scheme,jvals,nc", CASES)
def test_rotsym_lines_match_oracle(group, scheme, jvals, nc, mode,
                                   precision):
    prob = _problem(group, scheme, jvals, nc)
    rows = prob.state_table()
    initials = list(range(0, len(rows), max(1, len(rows) // 40)))
    tr, _ = orc.run(prob, initials)
    ref = oracle_map(tr)
    with Engine.from_problem(prob) as eng:
        got = engine_lines(eng, prob, initials, mode, precision)
    assert len(ref) > 0
    if precision == 0:
        compare_lines(got, ref)
    else:
        assert set(got) == set(ref)
        assert {k for k in got if got[k]["emitted"]} == \
            {k for k in ref if ref[k]["emitted"]}
        r = np.array([ref[k]["S"] for k in ref])
        g = np.array([got[k]["S"] for k in ref])
        strong = r > 1e-6 * r.max()
        assert (np.abs(g - r)[strong] / r[strong]).max() <= 1e-7


def test_rotsym_single_half_ls_matches_literal_kernel():
    """gain_half_ls on a symmetric-top pair against the oracle's literal
    ROTSYM kernel (dipole_GPU.cu:641-723)."""
    prob = _problem("C3V(M)", "c1", [3, 4], 5)
    bI, bF = prob.bases
    rng = np.random.default_rng(3)
    c = rng.standard_normal(bI.ntot[0])
    x = orc.expand(bI, 0, 0, c)
    ref = orc.half_ls_rotsym(bI.nlevels, bF.nlevels, bI.rmd, bI.deg, bF.deg,
                             bI.level, bF.level, bI.vib, bF.vib, 0, prob.nc,
                             prob.mu, prob.nc, bI.wig[2], x)
    with Engine.from_problem(prob) as eng:
        h = eng.half_ls(0, 0, 0, c, 1, bF.nprim)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(h, ref, rtol=0, atol=1e-13 * scale)


@pytest.mark.parametrize("precision", [0, 8])
def test_rotsym_sweep_matches_unbatched(precision):
    """Symmetric-top tasks in a captured sweep (their gathers run as
    launches of their own inside the graph) against the plain path."""
    import torch
    prob = _problem("TD(M)", "td5", [2, 3, 4], 10)
    rng = np.random.default_rng(9)
    with Engine.from_problem(prob) as eng:
        tasks, outs = [], []
        pairs = {g: eng.gamma_pair(g) for g in range(prob.params.nrepres)}
        for (jI, gI), blk in sorted(prob.blocks.items()):
            levels = sorted(rng.choice(blk.n, size=min(4, blk.n),
                                       replace=False).tolist())
            for (jF, gF), bF in sorted(prob.blocks.items()):
                if gF != pairs[gI] or \
                        abs(prob.bases[jF].J - prob.bases[jI].J) > 1:
                    continue
                o = {"S": torch.zeros((len(levels), bF.n),
                                      dtype=torch.float64, device="cuda:0"),
                     "flags": torch.zeros((len(levels), bF.n),
                                          dtype=torch.uint8,
                                          device="cuda:0")}
                tasks.append((jI, gI, levels, jF, gF,
                              {k: v.data_ptr() for k, v in o.items()}))
                outs.append(o)
        sw = Sweep(eng, tasks, precision=precision)
        for _ in range(2):
            sw.run()
            eng.synchronize()
            n = 0
            for (jI, gI, levels, jF, gF, _), o in zip(tasks, outs):
                fl, S = o["flags"].cpu().numpy(), o["S"].cpu().numpy()
                for i, lv in enumerate(levels):
                    r = eng.plain_lines(jI, gI, lv, jF, gF, want_ls=False)
                    assert np.array_equal(r["flags"], fl[i])
                    ref = r["S"]
                    if ref.max() <= 0:
                        continue
                    tol = 1e-12 if precision == 0 else 1e-7
                    strong = ref > 1e-6 * ref.max()
                    assert (np.abs(S[i] - ref)[strong]
                            <= tol * ref[strong]).all()
                    n += int((fl[i] & 1).sum())
            assert n > 0
        sw.close()


def test_rotsym_writer_matches_oracle(tmp_path):
    prob = _problem("C3V(M)", "c3v3", [2, 3, 4], 6)
    prob.params.erange_lower = (0.0, 6000.0)
    rows = prob.state_table()
    initials = [i for i, r in enumerate(rows) if r[0] <= 6000.0]
    tr, _ = orc.run(prob, initials)
    em = tr[tr["emitted"] != 0]
    em = em[np.lexsort((em["iF"], em["iI"]))]
    path = tmp_path / "lines.out"
    with Engine.from_problem(prob) as eng:
        n = eng.write_transitions(path, full_line=False, intensity=True,
                                  batch=6)
    lines = open(path).read().splitlines()
    assert n == len(lines) == len(em) > 0
    for t, line in zip(em, lines):
        eI, jindI, gI, lI = rows[t["iI"]]
        eF, jindF, gF, lF = rows[t["iF"]]
        head, _, rest = line.partition("<-")
        assert head.split()[0] == ("%12.6f" % (eF - eI)).strip()
        v = rest.split()
        assert float(v[3]) == pytest.approx(t["A"], rel=2e-9)
        assert float(v[4]) == pytest.approx(t["I"], rel=2e-9)


def test_rotsym_noalloc_and_errors():
    """The allocation-free plain call on symmetric-top bases equals
    gain_plain_lines bit for bit; the adjoint and a pair without its
    rotational table are refused."""
    import torch
    prob = _problem("C3V(M)", "c1", [2, 3], 6)
    with Engine.from_problem(prob) as eng:
        L = _lib.load()
        nF = prob.blocks[(1, 0)].n
        outs = []
        for _ in range(2):
            d = {k: torch.zeros(nF, dtype=torch.float64, device="cuda:0")
                 for k in ("S", "A", "I")}
            d["ls"] = torch.zeros(nF, dtype=torch.float64, device="cuda:0")
            d["flags"] = torch.zeros(nF, dtype=torch.uint8, device="cuda:0")
            outs.append(d)

        def lines(d):
            s = _lib.Lines()
            for k in ("ls", "S", "A", "I", "flags"):
                setattr(s, k, d[k].data_ptr())
            s.on_device = 1
            return s
        torch.cuda.synchronize()
        _lib.check(L.gain_plain_lines_noalloc(eng._h, 0, 0, 2, None, 1, 0,
                                              C.byref(lines(outs[0]))))
        _lib.check(L.gain_plain_lines(eng._h, 0, 0, 2, None, 1, 0,
                                      C.byref(lines(outs[1]))))
        eng.synchronize()
        for k in outs[0]:
            assert torch.equal(outs[0][k], outs[1][k]), k
        assert int((outs[0]["flags"] & 1).sum()) > 0
        dS = torch.ones(nF, dtype=torch.float64, device="cuda:0")
        dmu = torch.zeros(3 * prob.nc ** 2, dtype=torch.float64,
                          device="cuda:0")
        with pytest.raises(_lib.GainError) as e:
            eng.adjoint(0, 0, [2], 1, 0, dS.data_ptr(), dmu.data_ptr())
        assert e.value.code == _lib.ERR_ARG
    # without the J=3 -> J=2 table the pair is refused
    prob2 = _problem("C3V(M)", "c1", [2, 3], 6)
    prob2.bases[1].wig = {k: v for k, v in prob2.bases[1].wig.items()
                          if k != 0}
    with Engine.from_problem(prob2) as eng:
        with pytest.raises(_lib.GainError) as e:
            eng.plain_lines(1, 0, 0, 0, 0)
        assert e.value.code == _lib.ERR_STATE
