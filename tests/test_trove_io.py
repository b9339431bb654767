"""TROVE checkpoint readers (host code, no GPU): dipole records (incl.
subrecords), contracted bases with the J=0 correlation, state descriptions
with the state filter, eigenvector records, and the error paths.  The files
are written by tests/trove_files.py in the formats the reference reads."""
import os
import struct

import numpy as np
import pytest

from gain_engine import _lib, synth, trove

import trove_files as tf


@pytest.mark.parametrize("max_sub", [2 ** 31 - 9, 1000, 7])
def test_dipole_round_trip_and_subrecords(tmp_path, max_sub):
    prob = synth.make_problem([1], 12, "c1")
    p = tmp_path / "j0_extfield.chk"
    tf.write_dipole(p, prob.mu, max_sub)
    assert np.array_equal(trove.read_dipole(p), prob.mu)


def test_dipole_orientation_from_hand_made_file(tmp_path):
    """One nonzero element m(i=1, f=2) of component 3 in the file's
    column-major matrix lands at piece[2][1][2] (TroveDipole.cpp:110: piece
    index f + i*ncontrF + k*ncontr*ncontrF from temp[i + f*ncontr + ...])."""
    nc = 3
    m = np.zeros(nc * nc)
    m[1 + 2 * nc] = 7.5
    p = tmp_path / "j0_extfield.chk"
    with open(p, "wb") as f:
        def rec(b):
            f.write(struct.pack("<i", len(b)) + b + struct.pack("<i", len(b)))
        rec(b"Start external field")
        rec(struct.pack("<i", nc))
        for a in range(3):
            rec(struct.pack("<i", a + 1))
            rec((m if a == 2 else np.zeros(nc * nc)).tobytes())
        rec(b"End external field  ")
    mu = trove.read_dipole(p)
    assert mu[2, 1, 2] == 7.5
    assert np.count_nonzero(mu) == 1


def test_dipole_errors(tmp_path):
    prob = synth.make_problem([1], 4, "c1")
    p = tmp_path / "d.chk"
    tf.write_dipole(p, prob.mu, footer=b"Bogus footer here!")
    with pytest.raises(_lib.GainError, match="footer"):
        trove.read_dipole(p)
    data = open(tmp_path / "d.chk", "rb").read()
    open(tmp_path / "t.chk", "wb").write(data[:100])
    with pytest.raises(_lib.GainError, match="truncated"):
        trove.read_dipole(tmp_path / "t.chk")
    with pytest.raises(_lib.GainError, match="not found"):
        trove.read_dipole(tmp_path / "missing.chk")


@pytest.mark.parametrize("scheme,jvals,nc", [
    ("c1", [0, 1, 3], 5), ("c2v4", [2, 5], 8), ("c3v3", [1, 2], 6),
    ("pair2", [4], 4)])
def test_basis_round_trip(tmp_path, scheme, jvals, nc):
    prob = synth.make_problem(jvals, nc, scheme)
    tf.write_problem(tmp_path, prob, scheme)
    for b in prob.bases:
        r = trove.read_basis(tmp_path / ("j0eigen_quanta%d.chk" % b.J),
                             tmp_path / "j0eigen_quanta0.chk", b.J,
                             prob.params.sym_degen)
        for k in ("kstart", "kblock", "tau", "vib", "icontr", "ijterms"):
            assert np.array_equal(r[k], getattr(b, k)), k
        assert r["ntot"] == list(b.ntot) and r["matsize"] == b.matsize
        for g in range(len(b.ntot)):
            assert np.array_equal(r["N"][g], b.N[g])
            assert np.array_equal(r["repres"][g], b.repres[g])


def test_correlation_matches_by_content_not_position(tmp_path):
    """Cases listed in reverse vibrational order inside every rotational
    level: the vibrational index of each function still comes from the J=0
    function with the same class quanta (TroveBasisSet.cpp:331-396)."""
    gs, nc, J = 4, 8, 2
    b = synth.make_basis(J, nc, "c2v4")
    tf.write_quanta(tmp_path / "q2.chk", b, gs, reverse_cases=True)
    tf.write_quanta(tmp_path / "q0.chk", synth.make_basis(0, nc, "c2v4"), gs)
    r = trove.read_basis(tmp_path / "q2.chk", tmp_path / "q0.chk", J,
                         [1, 1, 1, 1])
    # function p of the file is member lam of group q, groups reversed
    per = nc // gs
    exp = []
    for slot in range(2 * J + 1):
        for j in range(per):
            q = per - 1 - j
            exp.extend(q * gs + lam for lam in range(gs))
    assert np.array_equal(r["vib"], exp)
    assert np.array_equal(r["kblock"], b.kblock)


def test_basis_errors(tmp_path):
    b = synth.make_basis(1, 4, "c1")
    tf.write_quanta(tmp_path / "q1.chk", b, 1)
    tf.write_quanta(tmp_path / "q0.chk", synth.make_basis(0, 4, "c1"), 1)
    txt = open(tmp_path / "q1.chk").read()
    open(tmp_path / "bad1.chk", "w").write(txt.replace(
        "Start Primitive basis set", "Start Something else"))
    with pytest.raises(_lib.GainError, match="bad header"):
        trove.read_basis(tmp_path / "bad1.chk", tmp_path / "q0.chk", 1, [1])
    open(tmp_path / "bad2.chk", "w").write(txt.replace(
        "End irreducible transformation", "End"))
    with pytest.raises(_lib.GainError, match="irrep-footer"):
        trove.read_basis(tmp_path / "bad2.chk", tmp_path / "q0.chk", 1, [1])
    # a J=0 basis without the vibrational functions of J=1
    tf.write_quanta(tmp_path / "q0small.chk", synth.make_basis(0, 2, "c1"), 1)
    with pytest.raises(_lib.GainError, match="no J=0 counterpart"):
        trove.read_basis(tmp_path / "q1.chk", tmp_path / "q0small.chk", 1,
                         [1])


def test_states_filter_labels_and_records(tmp_path):
    prob = synth.make_problem([2, 3], 8, "c2v4")
    tf.write_problem(tmp_path, prob, "c2v4", extra_high=(40000.0, 50000.0),
                     label_of=lambda s: 2 * s + 3, shuffle_seed=5)
    for (jind, g), blk in prob.blocks.items():
        J = prob.bases[jind].J
        s = trove.read_states(tmp_path / ("j0eigen_descr%d_%d.chk" %
                                          (J, g + 1)), g, prob.params)
        assert s["rec_len"] == blk.rec_len
        # the two states above every window are dropped (States.cpp:41-53)
        assert np.array_equal(s["energy"], blk.energy)
        assert np.array_equal(s["ilevel"], 2 * np.arange(blk.n) + 3)
        recs = trove.read_records(
            tmp_path / ("j0eigen_vectors%d_%d.chk" % (J, g + 1)), s["irec"],
            s["rec_len"])
        assert np.array_equal(recs, blk.vectors)


def test_states_errors(tmp_path):
    prob = synth.make_problem([1], 4, "c1")
    tf.write_problem(tmp_path, prob, "c1")
    p = tmp_path / "j0eigen_descr1_1.chk"
    txt = open(p).read()
    open(tmp_path / "nofoot.chk", "w").write(
        txt.replace("End Quantum numbers and energies", ""))
    with pytest.raises(_lib.GainError, match="not complete"):
        trove.read_states(tmp_path / "nofoot.chk", 0, prob.params)
    q = synth.Params(**{**prob.params.__dict__, "nrepres": 2,
                        "sym_degen": [1, 1], "gns": [1.0, 1.0],
                        "isym_pairs": [1, 1], "isym_do": [1, 1],
                        "igamma_pair": [1, 0]})
    with pytest.raises(_lib.GainError, match="irrep 1 in the file of irrep 2"):
        trove.read_states(p, 1, q)
    with pytest.raises(_lib.GainError, match="out of the file"):
        trove.read_records(tmp_path / "j0eigen_vectors1_1.chk", [10 ** 6], 3)
