"""Real-data drop-in path: a problem written as TROVE checkpoint files
(tests/trove_files.py), loaded by gain_trove_load (dipole, bases with the
J=0 correlation, filtered states, shuffled records, level numbers) and run
through gain_write_transitions, gives byte-identical output to the same
problem uploaded directly; the line count matches the oracle."""
import numpy as np
import pytest

from gain_engine import synth
from gain_engine.engine import Engine
from oracle import oracle as orc

import trove_files as tf

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme,jvals,nc", [
    ("c2v4", [2, 3, 4], 8), ("c3v3", [1, 2, 3], 6), ("c1", [0, 1], 7)])
def test_trove_load_matches_direct_upload(tmp_path, scheme, jvals, nc):
    prob = synth.make_problem(jvals, nc, scheme)
    prob.params.erange_lower = (0.0, 5000.0)

    def label(s):
        return 3 * s + 1

    tf.write_problem(tmp_path, prob, scheme, max_sub=4096,
                     extra_high=(60000.0,), label_of=label, shuffle_seed=3)
    a = Engine(0, prob.jmax, prob.params.nrepres, prob.params.maxdeg)
    try:
        a.set_params(prob.params)
        a.trove_load(tmp_path, jvals)
        na = a.write_transitions(tmp_path / "trove.out", full_line=True)
    finally:
        a.close()
    with Engine.from_problem(prob) as b:
        for (jind, g), blk in prob.blocks.items():
            lv = np.array([label(s) for s in range(blk.n)], np.int32)
            b.set_block_levels(jind, g, lv)
        nb = b.write_transitions(tmp_path / "direct.out", full_line=True)
    ta = open(tmp_path / "trove.out").read()
    tb = open(tmp_path / "direct.out").read()
    assert na == nb > 0
    assert ta == tb
    rows = prob.state_table()
    init = [i for i, r in enumerate(rows) if 0.0 <= r[0] <= 5000.0]
    tr, _ = orc.run(prob, init)
    assert int((tr["emitted"] != 0).sum()) == na
