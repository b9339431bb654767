"""Td reduced symmetry and the filter inputs the other parity tests leave at
their defaults, on the GPU against the oracle.

* Td (symmetry_type 3): irreps A1, A2, E, F1, F2 with degeneracies
  1, 1, 2, 3, 3 (maxdeg 3, nine component pairs per F1 <-> F2 transition),
  selection pairs A1-A2, E-E, F1-F2 (TroveInput.cpp:202-236) and the
  reduced degeneracy rules of States.cpp:108-115 (F1 -> F2 from component
  1, F2 -> F1 from component 2, E -> E from component 1), and the same
  problem unreduced (every component pair).
* Filter inputs (States.cpp:46-76, Output.cpp:65,95): a zero-point energy
  that shifts both energy windows, a frequency window that starts below
  zero (pairs with nu < 1e-40 pass the filter but form no line), an
  intensity threshold on A, an irrep pair with gns = 0 and an irrep with
  isym_do = 0.

Plain FP64, batched FP64 and batched CRT (12 moduli) paths: transition
lists and emitted sets bit-exact, S, A, I within 1e-12 (FP64) or S within
1e-7 on strong lines (CRT)."""
import numpy as np
import pytest

from gain_engine import synth
from gain_engine.engine import Engine
from oracle import oracle as orc

from helpers import compare_lines, engine_lines, oracle_map

pytestmark = pytest.mark.gpu


def _td(reduced):
    prob = synth.make_problem([2, 3], 30, "td5")
    prob.params.reduced = reduced
    prob.params.erange_lower = (0.0, 1e9)
    return prob


def _edge_problem():
    prob = synth.make_problem([2, 3, 4], 20, "td5")
    p = prob.params
    p.zpe = 450.0
    p.gns = [0.0, 0.0, 2.0, 3.0, 3.0]   # the A1-A2 pair carries no lines
    p.isym_do = [1, 1, 0, 1, 1]         # E states neither lower nor upper
    p.erange_lower = (100.0, 7000.0)    # on E - ZPE
    p.erange_upper = (300.0, 9500.0)
    p.freq_window = (-800.0, 7000.0)
    rows = prob.state_table()
    initials = list(range(0, len(rows), 2))
    # threshold on A between the 30th and 31st percentile of the lines
    tr, _ = orc.run(prob, initials)
    a = np.sort(tr["A"][tr["A"] > 0])
    p.threshold = float(0.5 * (a[int(0.3 * len(a))] +
                               a[int(0.3 * len(a)) + 1]))
    return prob, initials


def _modes():
    return [("plain", 0), ("batched", 0), ("batched", 8)]


def _check(prob, initials, mode, precision):
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
        assert strong.any()
        assert (np.abs(g - r)[strong] / r[strong]).max() <= 1e-7
    return tr, ref


@pytest.mark.parametrize("mode,precision", _modes())
@pytest.mark.parametrize("reduced", [True, False])
def test_td_every_direction(reduced, mode, precision):
    prob = _td(reduced)
    rows = prob.state_table()
    initials = list(range(0, len(rows), max(1, len(rows) // 60)))
    tr, ref = _check(prob, initials, mode, precision)
    # every selection pair of Td carries lines, including F1 <-> F2
    pairs = {(rows[t["iI"]][2], rows[t["iF"]][2]) for t in tr}
    assert {(0, 1), (1, 0), (2, 2), (3, 4), (4, 3)} <= pairs
    if not reduced:
        # unreduced F1 -> F2 lines use all nine component pairs
        t = next(t for t in tr if rows[t["iI"]][2] == 3 and t["S"] > 0)
        assert np.count_nonzero(t["ls"]) == 9


@pytest.mark.parametrize("mode,precision", _modes())
def test_filter_edge_inputs(mode, precision):
    prob, initials = _edge_problem()
    rows = prob.state_table()
    tr, ref = _check(prob, initials, mode, precision)
    p = prob.params
    gI = np.array([rows[t["iI"]][2] for t in tr])
    gF = np.array([rows[t["iF"]][2] for t in tr])
    nu = np.array([rows[t["iF"]][0] - rows[t["iI"]][0] for t in tr])
    # the inputs are exercised: no line of a gns = 0 or isym_do = 0 irrep;
    # pairs below nu = 0 pass the filter without a line; the threshold
    # drops some lines that pass the filter
    assert not np.isin(gI, [0, 1, 2]).any() and not np.isin(gF, [0, 1, 2]).any()
    assert (nu < 0).any() and (tr["emitted"][nu < 0] == 0).all()
    dropped = (nu > 0) & (tr["emitted"] == 0)
    assert dropped.any() and (tr["A"][dropped] < p.threshold).all()
    assert (tr["emitted"] != 0).any()
    eI = np.array([rows[t["iI"]][0] for t in tr]) - p.zpe
    assert eI.min() >= p.erange_lower[0]


@pytest.mark.parametrize("case", ["td", "edge"])
def test_write_transitions_td_and_edge_inputs(tmp_path, case):
    """The whole emitted list through gain_write_transitions (integer
    fields and frequency exact, A and intensity within the printed digits,
    components in the full-line form)."""
    if case == "td":
        prob = _td(True)
    else:
        prob, _ = _edge_problem()
    p = prob.params
    rows = prob.state_table()
    initials = [i for i, r in enumerate(rows)
                if p.erange_lower[0] <= r[0] - p.zpe <= p.erange_lower[1]
                and p.isym_do[r[2]]]
    tr, _ = orc.run(prob, initials)
    em = tr[tr["emitted"] != 0]
    em = em[np.lexsort((em["iF"], em["iI"]))]
    path = tmp_path / "lines.out"
    with Engine.from_problem(prob) as eng:
        n = eng.write_transitions(path, full_line=True, intensity=True,
                                  batch=5)
    lines = open(path).read().splitlines()
    assert n == len(lines) == len(em) > 0
    ps = orc.params_struct(p)
    for t, line in zip(em, lines):
        eI, jindI, gI, lI = rows[t["iI"]]
        eF, jindF, gF, lF = rows[t["iF"]]
        head, _, rest = line.partition("<-")
        f = head.split()
        assert f[0] == ("%12.6f" % (eF - eI)).strip()
        assert [int(v) for v in f[1:]] == [lF + 1, prob.bases[jindF].J,
                                           gF + 1]
        vals, _, comp = rest.partition("[")
        v = vals.split()
        assert [int(x) for x in v[:3]] == [lI + 1, prob.bases[jindI].J,
                                           gI + 1]
        assert float(v[3]) == pytest.approx(t["A"], rel=2e-9)
        assert float(v[4]) == pytest.approx(t["I"], rel=2e-9)
        nd = p.sym_degen
        ref = [t["ls"][dI + dF * p.maxdeg]
               for dF in range(nd[gF]) for dI in range(nd[gI])
               if orc.lib().orc_degeneracy_filter(ps, gI, gF, dI, dF)]
        got = [float(x) for x in comp.split("]")[0].split()]
        np.testing.assert_allclose(got, ref, rtol=2e-9,
                                   atol=1e-12 * max(map(abs, ref)))
