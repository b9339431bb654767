"""Whole transition list (main.cpp:238-338 + Output.cpp:52-112) through
gain_write_transitions, against the oracle on multi-J problems: same set of
lines, same order (initial states by energy, finals by energy), integer
fields and the frequency bit-exact (both print the same FP64 difference),
A, intensity and components within the 10 printed digits."""
import numpy as np
import pytest

from gain_engine import synth
from gain_engine.engine import Engine
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _parse(path):
    rows = []
    with open(path) as f:
        for line in f:
            assert line.endswith(" || \n")
            head, _, rest = line.partition("<-")
            nu, lF, jF, gF = head.split()
            fields = rest.split("||")[0]
            ls = None
            if "[" in fields:
                fields, _, lsp = fields.partition("[")
                ls = [float(v) for v in lsp.split("]")[0].split()]
            vals = fields.split()
            rows.append((nu, int(lF) - 1, int(jF), int(gF) - 1,
                         int(vals[0]) - 1, int(vals[1]), int(vals[2]) - 1,
                         [float(v) for v in vals[3:]], ls))
    return rows


@pytest.mark.parametrize("scheme,jvals,nc,full", [
    ("c2v4", [2, 3, 4], 16, False), ("c3v3", [1, 2, 3], 12, True),
    ("c1", [0, 1, 2], 10, True)])
def test_write_transitions_matches_oracle(tmp_path, scheme, jvals, nc, full):
    prob = synth.make_problem(jvals, nc, scheme)
    prob.params.erange_lower = (0.0, 6000.0)
    rows = prob.state_table()
    initials = [i for i, r in enumerate(rows)
                if 0.0 <= r[0] - prob.params.zpe <= 6000.0]
    tr, _ = orc.run(prob, initials)
    em = tr[tr["emitted"] != 0]
    # reference order: initial by energy rank, then final by energy rank
    em = em[np.lexsort((em["iF"], em["iI"]))]
    path = tmp_path / "lines.out"
    with Engine.from_problem(prob) as eng:
        n = eng.write_transitions(path, full_line=full, intensity=True,
                                  batch=7)
    got = _parse(path)
    assert n == len(got) == len(em)
    for t, g in zip(em, got):
        eI, jindI, gI, lI = rows[t["iI"]]
        eF, jindF, gF, lF = rows[t["iF"]]
        assert g[0] == ("%12.6f" % (eF - eI)).strip()
        assert g[1:7] == (lF, prob.bases[jindF].J, gF, lI,
                          prob.bases[jindI].J, gI)
        A, inten = g[7]
        assert A == pytest.approx(t["A"], rel=2e-9)
        assert inten == pytest.approx(t["I"], rel=2e-9)
        if full:
            nd = prob.params.sym_degen
            ps = orc.params_struct(prob.params)
            ref = [t["ls"][dI + dF * prob.params.maxdeg]
                   for dF in range(nd[gF]) for dI in range(nd[gI])
                   if orc.lib().orc_degeneracy_filter(ps, gI, gF, dI, dF)]
            np.testing.assert_allclose(g[8], ref, rtol=2e-9,
                                       atol=1e-12 * max(map(abs, ref)))
