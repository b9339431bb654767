"""Pins the CPU oracle (CPU only, no GPU).

Each oracle function is checked against something independent of its own
formula: exact third-party 3j values, the Fortran textbook form of the
rotational factor, dense brute-force operators, exactly rounded sums,
orthogonality / isometry / sum-rule invariants, hand-enumerated filter
cases and literal closed-form numbers."""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from gain_engine import synth
from oracle import oracle as orc

from helpers import dense_half_ls, irregular_basis, textbook_rot

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- 3j table
def test_threej_table_matches_sympy_exact_values():
    data = json.load(open(os.path.join(GOLDEN, "threej_jmax31.json")))
    jmax = data["jmax"]
    tab = orc.threej_table(jmax)
    n = 0
    for jI, kI, jF, kF, v in data["entries"]:
        ref = float(v)
        got = tab[orc.threej_index(jmax, jI, kI, jF, kF)]
        assert abs(got - ref) <= 4e-16 * max(abs(ref), 1e-300) + 1e-300, (
            jI, kI, jF, kF, got, ref)
        n += 1
    assert n == len(data["entries"]) > 4000
    # every table slot not listed stays zero (GpuManager.cpp:159-162)
    listed = {orc.threej_index(jmax, e[0], e[1], e[2], e[3])
              for e in data["entries"]}
    for i in range(tab.shape[0]):
        if i not in listed:
            assert tab[i] == 0.0


@pytest.mark.parametrize("j1", [1, 2, 5, 9])
def test_threej_orthogonality(j1):
    # sum_{m1,m2} (2 j3 + 1) (j1 1 j3; m1 m2 m3)(j1 1 j3'; m1 m2 m3) = delta
    for j3 in range(j1 - 1, j1 + 2):
        for j3p in range(j1 - 1, j1 + 2):
            for m3 in range(-min(j3, j3p), min(j3, j3p) + 1):
                s = 0.0
                for m2 in (-1, 0, 1):
                    m1 = -m2 - m3
                    if abs(m1) > j1:
                        continue
                    s += orc.three_j(j1, 1, j3, m1, m2, m3) * \
                        orc.three_j(j1, 1, j3p, m1, m2, m3)
                s *= 2 * j3 + 1
                assert s == pytest.approx(1.0 if j3 == j3p else 0.0,
                                          abs=1e-14)


def test_threej_closed_form_j2_1():
    # (j j 1; m -m 0) = (-1)^(j-m) m / sqrt(j (j+1) (2j+1))  (Edmonds
    # Table 2); the odd column permutation to (j 1 j; m 0 -m) flips the sign
    for j in range(1, 20):
        for m in range(-j, j + 1):
            ref = (-1) ** (j - m) * m / math.sqrt(j * (j + 1) * (2 * j + 1))
            assert orc.three_j(j, j, 1, m, -m, 0) == pytest.approx(
                ref, abs=1e-15)
            assert orc.three_j(j, 1, j, m, 0, -m) == pytest.approx(
                -ref, abs=1e-15)


# ---------------------------------------------------------------- rot factor
def test_rot_coeff_matches_fortran_textbook_form():
    jmax = 12
    tab = orc.threej_table(jmax)
    trunc = 0.70710678118 * math.sqrt(2.0)  # Q1: truncated 1/sqrt(2)
    checked = 0
    for jI in range(0, jmax):
        for jF in range(max(jI - 1, 0), jI + 2):
            if jF > jmax:
                continue
            for kI in range(jI + 1):
                for kF in range(max(kI - 1, 0), min(kI + 1, jF) + 1):
                    f3j = tab[orc.threej_index(jmax, jI, kI, jF, kF)]
                    for tI in (0, 1):
                        for tF in (0, 1):
                            v, a = orc.rot_coeff(jmax, tab, jI, jF, kI, tI,
                                                 kF, tF)
                            ta, tv = textbook_rot(jI, jF, kI, tI, kF, tF, f3j)
                            assert a == ta
                            if kI != kF and kI * kF != 0:
                                assert v == pytest.approx(tv * trunc,
                                                          rel=1e-15,
                                                          abs=1e-300)
                            else:
                                assert v == tv
                            checked += 1
    assert checked > 1000


# ---------------------------------------------------------------- expand
def test_expand_identity_and_hadamard_isometry():
    rng = np.random.default_rng(7)
    b1 = synth.make_basis(2, 8, "c1")
    c = rng.standard_normal(b1.ntot[0])
    assert np.array_equal(orc.expand(b1, 0, 0, c), c)
    b4 = synth.make_basis(3, 8, "c2v4")
    for g in range(4):
        c = rng.standard_normal(b4.ntot[g])
        x = orc.expand(b4, g, 0, c)
        assert np.linalg.norm(x) == pytest.approx(np.linalg.norm(c),
                                                  rel=1e-14)


def _dense_P(b, g, d):
    """P[r, t] assembled from the definition of the expansion."""
    P = np.zeros((b.nprim, b.ntot[g]))
    ijt = b.ijterms[g * b.maxsymcoeffs:(g + 1) * b.maxsymcoeffs]
    for r in range(b.nprim):
        case, lam = b.icontr[r], b.icontr[r + b.nprim]
        for i in range(b.N[g][case]):
            t = ijt[case] + i
            P[r, t] = b.repres[g][t + d * b.ntot[g] +
                                  lam * b.sdeg[g] * b.ntot[g]]
    return P


@pytest.mark.parametrize("scheme", ["c2v4", "c3v3", "pair2", "td5"])
def test_expand_matches_dense_operator(scheme):
    rng = np.random.default_rng(11)
    nc = {"c2v4": 8, "c3v3": 6, "pair2": 6, "td5": 10}[scheme]
    b = synth.make_basis(2, nc, scheme)
    for g in range(len(b.ntot)):
        if b.ntot[g] == 0:
            continue
        for d in range(b.sdeg[g]):
            P = _dense_P(b, g, d)
            c = rng.standard_normal(b.ntot[g])
            np.testing.assert_allclose(orc.expand(b, g, d, c), P @ c,
                                       rtol=0, atol=1e-14)
            # every component is an isometry for these generators
            np.testing.assert_allclose(P.T @ P, np.eye(b.ntot[g]),
                                       atol=1e-14)
    # the components of all irreps together are an orthogonal change of
    # basis of the primitives
    cols = [_dense_P(b, g, d) for g in range(len(b.ntot))
            for d in range(b.sdeg[g]) if b.ntot[g]]
    Pall = np.hstack(cols)
    if Pall.shape[1] == b.nprim:
        np.testing.assert_allclose(Pall.T @ Pall, np.eye(b.nprim),
                                   atol=1e-14)


# ---------------------------------------------------------------- half ls
@pytest.mark.parametrize("jI,jF,seed", [(0, 1, 1), (2, 3, 2), (3, 3, 3),
                                        (3, 2, 4), (1, 0, 5)])
def test_half_ls_matches_dense_textbook_operator(jI, jF, seed):
    rng = np.random.default_rng(seed)
    nc = 6
    prob = synth.make_problem([jI, jF] if jI != jF else [jI], nc, "c1")
    bI = irregular_basis(jI, nc, rng, max_per_block=9)
    bF = irregular_basis(jF, nc, rng, max_per_block=9)
    x = rng.standard_normal(bI.nprim)
    jmax = max(jI, jF)
    prob.bases = [bI, bF]
    tab = orc.threej_table(jmax)
    h = orc.half_ls(prob, bI, bF, x, tab)
    ref = dense_half_ls(bI, bF, prob.mu, x, tab, jmax, orc.threej_index)
    scale = max(np.abs(ref).max(), 1e-300)
    np.testing.assert_allclose(h, ref, rtol=0, atol=1e-14 * scale)


def test_half_ls_dipole_piece_assigns_only_its_rows():
    # Q8: the kernel writes only entries whose vib index is in the piece
    rng = np.random.default_rng(3)
    nc = 8
    prob = synth.make_problem([2, 3], nc, "c1")
    bI, bF = prob.bases
    x = rng.standard_normal(bI.nprim)
    tab = orc.threej_table(3)
    full = orc.half_ls(prob, bI, bF, x, tab)
    part = orc.half_ls(prob, bI, bF, x, tab, startF=2, endF=6)
    inside = (bF.vib >= 2) & (bF.vib < 6)
    np.testing.assert_allclose(part[inside], full[inside], rtol=1e-15,
                               atol=1e-300)
    assert np.all(part[~inside] == 0.0)


# ---------------------------------------------------------------- ROTSYM
def _rotsym_levels(b):
    """Level index 2k + tau of every primitive of a K-block basis."""
    k = np.repeat(np.arange(b.J + 1), b.kblock)
    return 2 * k + (np.asarray(b.tau) & 1), 2 * (b.J + 1)


@pytest.mark.parametrize("jI,jF,seed,piece", [(2, 3, 1, (0, 7)),
                                              (3, 2, 2, (2, 6)),
                                              (3, 3, 3, (1, 7)),
                                              (1, 0, 4, (0, 7))])
def test_half_ls_rotsym_reduces_to_kblock_operator(jI, jF, seed, piece):
    """A symmetric-top basis written as ROTSYM levels (k, tau) with the
    Wigner table filled by the K-block rotational factor, scaled by an
    asymmetric degeneracy matrix c[dI, dF], is the K-block half line
    strength of x restricted to each initial degeneracy, mixed by c."""
    rng = np.random.default_rng(seed)
    nc = 7
    prob = synth.make_problem([jI, jF] if jI != jF else [jI], nc, "c1")
    bI = irregular_basis(jI, nc, rng, max_per_block=9)
    bF = irregular_basis(jF, nc, rng, max_per_block=9)
    prob.bases = [bI, bF]
    jmax = max(jI, jF)
    tab = orc.threej_table(jmax)
    lvI, nlI = _rotsym_levels(bI)
    lvF, nlF = _rotsym_levels(bF)
    c = np.array([[1.0, 2.0], [3.0, 5.0]])
    W = np.zeros((2, 2, nlF, nlI, 3))
    for kI in range(jI + 1):
        for tI in (0, 1):
            for kF in range(jF + 1):
                for tF in (0, 1):
                    v, a = orc.rot_coeff(jmax, tab, jI, jF, kI, tI, kF, tF)
                    if a >= 0:
                        for dI in (0, 1):
                            for dF in (0, 1):
                                W[dF, dI, 2 * kF + tF, 2 * kI + tI, a] = (
                                    v * c[dI, dF])
    degI = rng.integers(0, 2, bI.nprim)
    degF = rng.integers(0, 2, bF.nprim)
    x = rng.standard_normal(bI.nprim)
    s0, s1 = piece
    mu = np.ascontiguousarray(prob.mu[:, :, s0:s1])
    h = orc.half_ls_rotsym(nlI, nlF, 2, degI, degF, lvI, lvF, bI.vib,
                           bF.vib, s0, s1, mu, nc, W, x)
    parts = [orc.half_ls(prob, bI, bF, x * (degI == d), tab, s0, s1)
             for d in (0, 1)]
    ref = sum(c[d, degF] * parts[d] for d in (0, 1))
    scale = max(np.abs(ref).max(), 1e-300)
    np.testing.assert_allclose(h, ref, rtol=0, atol=1e-14 * scale)
    inside = (bF.vib >= s0) & (bF.vib < s1)
    assert np.all(h[~inside] == 0.0)


def test_half_ls_rotsym_delta_wigner_is_a_gathered_gemv():
    """W[n] = delta(n, z) leaves h = x . mu_z[vibI, vibF] (numpy GEMV);
    entries outside the dipole piece keep the caller's values."""
    rng = np.random.default_rng(11)
    nc, nI, nF, nlI, nlF, md = 9, 300, 200, 4, 6, 3
    mu = rng.standard_normal((3, nc, nc))
    vibI = rng.integers(0, nc, nI)
    vibF = rng.integers(0, nc, nF)
    degI, degF = rng.integers(0, md, nI), rng.integers(0, md, nF)
    lvI, lvF = rng.integers(0, nlI, nI), rng.integers(0, nlF, nF)
    x = rng.standard_normal(nI)
    s0, s1 = 2, 7
    sentinel = np.full(nF, 7.25)
    for z in range(3):
        W = np.zeros((md, md, nlF, nlI, 3))
        W[..., z] = 1.0
        h = orc.half_ls_rotsym(nlI, nlF, md, degI, degF, lvI, lvF, vibI,
                               vibF, s0, s1,
                               np.ascontiguousarray(mu[:, :, s0:s1]), nc, W,
                               x, out=sentinel)
        inside = (vibF >= s0) & (vibF < s1)
        ref = x @ mu[z][vibI][:, vibF]
        np.testing.assert_allclose(h[inside], ref[inside], rtol=0,
                                   atol=1e-13 * np.abs(ref).max())
        assert np.all(h[~inside] == 7.25)


# ---------------------------------------------------------------- dot
def test_dot_is_correctly_rounded_on_ill_conditioned_input():
    # condition number ~1e9: double-double accumulation still rounds to within
    # one ulp of the exact rational dot; a plain FP64 dot does not
    rng = np.random.default_rng(5)
    n = 2000
    x = rng.standard_normal(n) * np.exp(rng.uniform(-4, 4, n))
    y = rng.standard_normal(n)
    y[-1] = -math.fsum(x[:-1] * y[:-1]) / x[-1] + 1e-6
    exact = sum(Fraction(float(a)) * Fraction(float(b)) for a, b in zip(x, y))
    bound = abs(exact) * Fraction(1, 2**52)
    got = orc.dot(x, y)
    assert abs(Fraction(got) - exact) <= bound
    naive = 0.0
    for a, b in zip(x, y):
        naive += a * b
    assert abs(Fraction(naive) - exact) > bound


# ---------------------------------------------------------------- sum rule
def test_sum_rule_complete_orthonormal_finals():
    # sum_F ls_F^2 = ||P_F^T h||^2 when the final records are a complete
    # orthonormal basis of the irrep (J 0 -> 1, C1)
    nc = 20
    prob = synth.make_problem([0, 1], nc, "c1")
    p = prob.params
    p.freq_window = (-1e9, 1e9)
    p.erange_upper = (-1e9, 1e9)
    rows = prob.state_table()
    I = next(i for i, r in enumerate(rows) if r[1] == 0)
    tr, rows = orc.run(prob, [I])
    jind, g, lvl = rows[I][1], rows[I][2], rows[I][3]
    cI = prob.blocks[(jind, g)].vector(lvl)
    tab = orc.threej_table(prob.jmax)
    h = orc.half_ls(prob, prob.bases[0], prob.bases[1],
                    orc.expand(prob.bases[0], 0, 0, cI), tab)
    to_j1 = tr[[rows[f][1] == 1 for f in tr["iF"]]]
    assert len(to_j1) == prob.bases[1].nprim
    assert math.fsum(to_j1["ls"][:, 0] ** 2) == pytest.approx(
        math.fsum(h * h), rel=1e-13)


# ---------------------------------------------------------------- filters
def _params(**kw):
    p = synth.Params(nrepres=4, sym_degen=[1, 1, 1, 1], gns=[1, 1, 3, 3],
                     isym_pairs=[1, 1, 2, 2], maxJ=5,
                     erange_lower=(0.0, 1000.0), erange_upper=(0.0, 5000.0),
                     freq_window=(10.0, 2000.0))
    for k, v in kw.items():
        setattr(p, k, v)
    return orc.params_struct(p)


def test_gamma_pairs_and_errors():
    s = _params()
    s.igamma_pair[:4] = [9, 9, 9, 9]
    assert orc.lib().orc_gamma_pairs(s) == 0
    assert list(s.igamma_pair[:4]) == [1, 0, 3, 2]
    s = _params(isym_pairs=[1, 1, 1, 2])  # three irreps in one pair class
    assert orc.lib().orc_gamma_pairs(s) == -1
    s = _params(gns=[1, 2, 3, 3])  # pair with different gns
    assert orc.lib().orc_gamma_pairs(s) == -2


def test_filter_intensity_cases():
    s = _params()
    assert orc.lib().orc_gamma_pairs(s) == 0
    f = lambda *a: orc.lib().orc_filter_intensity(s, *a)
    assert f(1, 0, 100.0, 2, 1, 200.0) == 1
    assert f(1, 0, 100.0, 2, 1, 110.0) == 1  # nu == window low edge
    assert f(1, 0, 100.0, 2, 1, 109.999) == 0
    assert f(1, 0, 100.0, 2, 1, 2100.0) == 1  # nu == window high edge
    assert f(1, 0, 100.0, 2, 1, 2100.001) == 0
    assert f(1, 0, 100.0, 2, 2, 200.0) == 0  # not the pair irrep
    assert f(1, 0, 100.0, 3, 1, 200.0) == 0  # |dJ| > 1
    assert f(0, 0, 100.0, 0, 1, 200.0) == 0  # J_I + J_F < 1
    assert f(5, 0, 100.0, 5, 1, 200.0) == 0  # both at maxJ
    assert f(4, 0, 100.0, 5, 1, 200.0) == 1
    assert f(1, 0, 1000.5, 2, 1, 1500.0) == 0  # lower window
    assert f(1, 0, 100.0, 2, 1, 5000.5) == 0  # upper window (nu ok? no)
    s2 = _params(gns=[0, 0, 3, 3])
    orc.lib().orc_gamma_pairs(s2)
    assert orc.lib().orc_filter_intensity(s2, 1, 0, 100.0, 2, 1, 200.0) == 0
    s3 = _params()
    s3.isym_do[1] = 0
    orc.lib().orc_gamma_pairs(s3)
    assert orc.lib().orc_filter_intensity(s3, 1, 0, 100.0, 2, 1, 200.0) == 0


def test_degeneracy_filter_table():
    def mk(stype, red, degen):
        p = synth.Params(nrepres=len(degen), sym_degen=degen,
                         gns=[1] * len(degen),
                         isym_pairs=list(range(len(degen))), maxJ=3,
                         symmetry_type=stype, reduced=red)
        return orc.params_struct(p)
    L = orc.lib()
    s = mk(2, True, [1, 1, 2])
    assert L.orc_degeneracy_filter(s, 0, 2, 0, 1) == 1  # a 1-dim irrep
    got = {(dI, dF) for dI in range(2) for dF in range(2)
           if L.orc_degeneracy_filter(s, 2, 2, dI, dF)}
    assert got == {(1, 0)}
    s = mk(2, False, [1, 1, 2])  # not reduced: everything
    assert all(L.orc_degeneracy_filter(s, 2, 2, a, b)
               for a in range(2) for b in range(2))
    s = mk(3, True, [1, 1, 2, 3, 3])  # Td
    sel = {(gI, gF, dI, dF) for gI in range(5) for gF in range(5)
           for dI in range(3) for dF in range(3)
           if [1, 1, 2, 3, 3][gI] > 1 and [1, 1, 2, 3, 3][gF] > 1 and
           L.orc_degeneracy_filter(s, gI, gF, dI, dF)}
    assert sel == {(3, 4, 1, 0), (4, 3, 2, 0), (2, 2, 1, 0)}


# ---------------------------------------------------------------- S, A, I
def test_line_closed_form_numbers():
    s = _params(freq_window=(0.0, 1e9), erange_upper=(0.0, 1e9),
                temperature=296.0, partition=1000.0, threshold=0.0)
    S, A, I = (orc.C.c_double() for _ in range(3))
    ls = np.zeros(9)
    ls[0] = 1.0
    ok = orc.lib().orc_line(s, 0, 0, 1, 0.0, 1, 1, 1, 1000.0, ls,
                            orc.C.byref(S), orc.C.byref(A), orc.C.byref(I))
    assert ok == 1
    assert S.value == 1.0
    assert A.value == pytest.approx(313.618923, rel=1e-15)
    # the reference constants give c2 = h c / k = 1.43876724 cm K, within
    # 1e-5 of CODATA's 1.438776877 cm K
    c2 = 6.62606896e-27 * 2.99792458e+10 / 1.380658e-16
    assert c2 == pytest.approx(1.438776877, rel=1e-5)
    beta_nu = c2 * 1000.0 / 296.0
    ref_I = 1.327209365e-12 * 313.618923 * 1.0 * 3.0 * \
        (1.0 - math.exp(-beta_nu)) / (1000.0 ** 2 * 1000.0)
    assert I.value == pytest.approx(ref_I, rel=1e-14)
    assert I.value == pytest.approx(1.2479e-15, rel=1e-4)  # by hand
    # nu below 1e-40 is dropped; A below threshold is dropped
    assert orc.lib().orc_line(s, 0, 0, 1, 5.0, 1, 1, 1, 5.0, ls,
                              orc.C.byref(S), orc.C.byref(A),
                              orc.C.byref(I)) == 0
    s.threshold = 1e6
    assert orc.lib().orc_line(s, 0, 0, 1, 0.0, 1, 1, 1, 1000.0, ls,
                              orc.C.byref(S), orc.C.byref(A),
                              orc.C.byref(I)) == 0


def test_reduced_degenerate_strength_not_averaged():
    # Output.cpp:89-91: reduced, ndegI != 1, ndegF != 1 -> S = sum ls^2
    p = synth.Params(nrepres=3, sym_degen=[1, 1, 2], gns=[1, 1, 1],
                     isym_pairs=[1, 1, 2], maxJ=3, symmetry_type=2,
                     reduced=True, freq_window=(0, 1e9),
                     erange_upper=(0, 1e9))
    s = orc.params_struct(p)
    ls = np.zeros(9)
    ls[1 + 0 * 2] = 3.0  # (dI=1, dF=0): the only pair the filter keeps
    ls[0] = 100.0  # filtered out
    S, A, I = (orc.C.c_double() for _ in range(3))
    orc.lib().orc_line(s, 1, 2, 2, 0.0, 2, 2, 2, 100.0, ls, orc.C.byref(S),
                       orc.C.byref(A), orc.C.byref(I))
    assert S.value == 9.0
    s.reduced = 0  # not reduced: all four pairs, averaged over ndegI
    ls2 = np.array([1.0, 2.0, 3.0, 4.0, 0, 0, 0, 0, 0])
    orc.lib().orc_line(s, 1, 2, 2, 0.0, 2, 2, 2, 100.0, ls2, orc.C.byref(S),
                       orc.C.byref(A), orc.C.byref(I))
    assert S.value == (1 + 4 + 9 + 16) / 2.0


def test_format_line_literal():
    s = _params(freq_window=(0.0, 1e9), erange_upper=(0.0, 1e9))
    ls = np.zeros(9)
    ls[0] = 0.5
    buf = orc.C.create_string_buffer(1024)
    orc.lib().orc_format_line(s, buf, 1024, 0, 0, 1, 4, 0.0, 1, 1, 1, 9,
                              1234.5, ls, 1, 1)
    txt = buf.value.decode()
    A = 3.13618923e-7 * 1 * 0.25 * 1234.5 ** 3
    # "%12.6f %8d %4d %4d <- %8d %4d %4d %16.9E " (Output.cpp:99)
    assert txt.startswith(" 1234.500000       10    1    2 <-        5    0"
                          "    1 ")
    assert f"{A:16.9E} " in txt
    assert txt.endswith("[   5.000000000E-01 ]  || \n")


def test_plain_fp64_timing_mode_computes_the_same_lines():
    """The plain-FP64 accumulation mode (used only to time the CPU baseline)
    runs the same loop: same transitions, S within FP64 rounding."""
    prob = synth.make_problem([2, 3], 12, "c2v4")
    prob.params.erange_lower = (0.0, 1e9)
    init = list(range(0, 40, 3))
    dd, _ = orc.run(prob, init)
    st = {}
    pl, _ = orc.run(prob, init, plain_fp64=True, stats=st)
    assert st["seconds"] >= 0.0
    assert np.array_equal(dd[["iI", "iF", "emitted"]], pl[["iI", "iF",
                                                           "emitted"]])
    big = dd["S"] > 1e-6 * dd["S"].max()
    np.testing.assert_allclose(pl["S"][big], dd["S"][big], rtol=1e-10)
    # and the checker mode is back on afterwards
    again, _ = orc.run(prob, init)
    assert np.array_equal(again["S"], dd["S"])
