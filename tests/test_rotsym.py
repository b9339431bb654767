"""Symmetric-top (ROTSYM) rotational factors: the symmetrised rotational
functions and the contracted dipole table of the library
(gain_rotsym_functions / gain_rotsym_wigner, restating c_wignerGen,
trove_wrappers.f90:170-914, with the group tables of symmetry.f90), checked
on the CPU against what the mathematics fixes:

* the rotational levels of a symmetric top fall into the irreps the
  textbook K rules give (C3v(M), D3h(M)), level by level in count;
* in Td(M) and the abelian groups (C2h(M), Cs(EM), G4(M), G4(EM), D2h(M),
  D2h) the irrep counts are the reduction of D^(J) by the closed-form
  SO(3) characters, and in the abelian groups each level has one parity of
  k, matching its character under the rotation by pi about z;
* the functions are orthonormal and every group operation acts on the
  components of a level by the irrep's character (trace of each block);
* with no symmetry the table is the K-block rotational factor of the
  oracle (pinned against the Fortran textbook form in test_oracle.py);
* in any group, the ROTSYM half line strength (oracle of
  dipole_GPU.cu:641-723) with this table equals the K-block half line
  strength (oracle of dipole_GPU.cu:392-515) carried into the symmetrised
  rotational basis on both sides.
No GPU is used."""
import ctypes as C
import math

import numpy as np
import pytest

from gain_engine import _lib, synth
from oracle import oracle as orc

GROUPS = ["C(M)", "CS(M)", "C2V(M)", "C2H(M)", "CS(EM)", "G4(M)", "G4(EM)",
          "D2H(M)", "D2H", "C3V(M)", "D3H(M)", "TD(M)", "DNH(M)"]
NOPER = {"C(M)": 1, "CS(M)": 2, "C2V(M)": 4, "C2H(M)": 4, "CS(EM)": 4,
         "G4(M)": 4, "G4(EM)": 8, "D2H(M)": 8, "D2H": 8, "C3V(M)": 6,
         "D3H(M)": 12, "TD(M)": 24, "DNH(M)": 4}
# Abelian groups (symmetry.f90:140-364): characters [irrep][operation], the
# rotation angle of each operation's equivalent rotation (0 for the
# identity, pi for a twofold axis) and the operations equivalent to the
# twofold rotation about z.
KLEIN = [[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]]
D2H_M = [[1, 1, 1, 1, 1, 1, 1, 1], [1, 1, 1, 1, -1, -1, -1, -1],
         [1, 1, -1, -1, -1, -1, 1, 1], [1, 1, -1, -1, 1, 1, -1, -1],
         [1, -1, 1, -1, -1, 1, -1, 1], [1, -1, 1, -1, 1, -1, 1, -1],
         [1, -1, -1, 1, 1, -1, -1, 1], [1, -1, -1, 1, -1, 1, 1, -1]]
PI = math.pi
ABELIAN = {
    "C2H(M)": (KLEIN, [0, PI, PI, PI], [3]),
    "CS(EM)": ([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, 1, -1], [1, -1, -1, 1]],
               [0, PI, PI, PI], [1]),
    "G4(M)": (KLEIN, [0, PI, PI, PI], [3]),
    "G4(EM)": ([[1, 1, 1, 1, 1, 1, 1, 1], [1, 1, -1, -1, 1, 1, -1, -1],
                [1, -1, -1, 1, 1, -1, -1, 1], [1, -1, 1, -1, 1, -1, 1, -1],
                [1, 1, -1, -1, -1, -1, 1, 1], [1, 1, 1, 1, -1, -1, -1, -1],
                [1, -1, 1, -1, -1, 1, -1, 1], [1, -1, -1, 1, -1, 1, 1, -1]],
               [0, PI, PI, PI, PI, PI, PI, 0], [3, 4]),
    "D2H(M)": (D2H_M, [0, PI, PI, PI, PI, PI, PI, 0], [1, 6]),
    "D2H": ([[1, 1, 1, 1, 1, 1, 1, 1], [1, 1, 1, 1, -1, -1, -1, -1],
             [1, 1, -1, -1, 1, 1, -1, -1], [1, 1, -1, -1, -1, -1, 1, 1],
             [1, -1, 1, -1, 1, -1, 1, -1], [1, -1, 1, -1, -1, 1, -1, 1],
             [1, -1, -1, 1, 1, -1, -1, 1], [1, -1, -1, 1, -1, 1, 1, -1]],
            [0, PI, PI, PI, PI, 0, PI, PI], [1, 4]),
}
# Td(M): classes E, (123), (12)(34), (1423)*, (12)*; sizes; the rotation
# angle of the equivalent rotation of each class; characters of
# A1, A2, E, F1, F2
TD_SIZES = [1, 8, 3, 6, 6]
TD_ANGLE = [0.0, 2 * math.pi / 3, math.pi, math.pi / 2, math.pi]
TD_CHARS = [[1, 1, 1, 1, 1], [1, 1, 1, -1, -1], [2, -1, 2, 0, 0],
            [3, 0, -1, 1, -1], [3, 0, -1, -1, 1]]


def functions(group, J):
    L = _lib.load()
    nl, md = C.c_int(), C.c_int()
    _lib.check(L.gain_rotsym_functions(group.encode(), J, C.byref(nl),
                                       C.byref(md), None, None, None, None))
    n = 2 * J + 1
    irr = np.zeros(nl.value, np.int32)
    root = np.zeros(nl.value * md.value, np.int32)
    vec = np.zeros(n * n)
    ops = np.zeros(NOPER[group] * n * n)
    _lib.check(L.gain_rotsym_functions(
        group.encode(), J, C.byref(nl), C.byref(md), irr.ctypes.data,
        root.ctypes.data, vec.ctypes.data, ops.ctypes.data))
    return (nl.value, md.value, irr, root.reshape(nl.value, md.value),
            vec.reshape(n, n), ops.reshape(NOPER[group], n, n))


def wigner(group, jI, jF):
    L = _lib.load()
    a, b, m = C.c_int(), C.c_int(), C.c_int()
    _lib.check(L.gain_rotsym_wigner(group.encode(), jI, jF, C.byref(a),
                                    C.byref(b), C.byref(m), None))
    rot = np.zeros(3 * a.value * b.value * m.value * m.value)
    _lib.check(L.gain_rotsym_wigner(group.encode(), jI, jF, C.byref(a),
                                    C.byref(b), C.byref(m), rot.ctypes.data))
    # first index fastest: [degF][degI][levelF][levelI][n]
    return a.value, b.value, m.value, rot.reshape(m.value, m.value, b.value,
                                                  a.value, 3)


def textbook_levels(group, J):
    """Irrep counts of the 2J+1 rotational functions of a symmetric top by
    the K rules of the molecular symmetry group (Bunker & Jensen)."""
    if group == "C3V(M)":  # A1, A2, E
        c = [0, 0, 0]
        c[0 if J % 2 == 0 else 1] += 1
        for k in range(1, J + 1):
            if k % 3 == 0:
                c[0] += 1
                c[1] += 1
            else:
                c[2] += 1
        return c
    if group == "D3H(M)":  # A1', A2', E', A1", A2", E"
        c = [0] * 6
        c[0 if J % 2 == 0 else 1] += 1
        for k in range(1, J + 1):
            r = k % 6
            if r == 0:
                c[0] += 1
                c[1] += 1
            elif r in (1, 5):
                c[5] += 1
            elif r in (2, 4):
                c[2] += 1
            else:
                c[3] += 1
                c[4] += 1
        return c
    raise ValueError(group)


@pytest.mark.parametrize("J", [0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 13])
def test_td_levels_reduce_the_rotation_group_representation(J):
    """A spherical top's 2J+1 functions span D^(J); its reduction over the
    rotations equivalent to the Td(M) operations (character of a rotation by
    a: sin((J + 1/2) a) / sin(a/2)) fixes the irrep counts."""
    def chi(a):
        return 2 * J + 1 if a == 0.0 else \
            math.sin((J + 0.5) * a) / math.sin(0.5 * a)
    want = [round(sum(n * c[k] * chi(a) for k, (n, a) in
                      enumerate(zip(TD_SIZES, TD_ANGLE))) / 24)
            for c in TD_CHARS]
    nl, md, irr, root, vec, ops = functions("TD(M)", J)
    assert md == 3
    assert np.bincount(irr, minlength=5).tolist() == want


@pytest.mark.parametrize("group", ["C3V(M)", "D3H(M)"])
@pytest.mark.parametrize("J", [0, 1, 2, 3, 4, 5, 6, 7, 9, 12])
def test_levels_follow_the_symmetric_top_k_rules(group, J):
    nl, md, irr, root, vec, ops = functions(group, J)
    want = textbook_levels(group, J)
    got = np.bincount(irr, minlength=len(want)).tolist()
    assert got == want
    degen = [1, 1, 2] if group == "C3V(M)" else [1, 1, 2, 1, 1, 2]
    assert sum(degen[s] for s in irr) == 2 * J + 1


# character tables [irrep][operation] with operations class by class
CHARS = {
    "C(M)": [[1]],
    "CS(M)": [[1, 1], [1, -1]],
    "C2V(M)": [[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]],
    "C3V(M)": [[1, 1, 1, 1, 1, 1], [1, 1, 1, -1, -1, -1],
               [2, -1, -1, 0, 0, 0]],
    "D3H(M)": [[1] * 12, [1, 1, 1, -1, -1, -1, 1, 1, 1, -1, -1, -1],
               [2, -1, -1, 0, 0, 0, 2, -1, -1, 0, 0, 0],
               [1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, -1],
               [1, 1, 1, -1, -1, -1, -1, -1, -1, 1, 1, 1],
               [2, -1, -1, 0, 0, 0, -2, 1, 1, 0, 0, 0]],
    "TD(M)": [sum(([c[k]] * n for k, n in enumerate(TD_SIZES)), [])
              for c in TD_CHARS],
}
CHARS.update({g: v[0] for g, v in ABELIAN.items()})
# DNH(M) at order 1 (E, sigma_v, i, C'2): A1', A2', A1", A2"
CHARS["DNH(M)"] = [[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1],
                   [1, -1, -1, 1]]


def so3_character(J, a):
    """Character of D^(J) for a rotation by a: sin((J + 1/2) a) / sin(a/2)."""
    return 2 * J + 1 if a == 0.0 else \
        math.sin((J + 0.5) * a) / math.sin(0.5 * a)


@pytest.mark.parametrize("group", sorted(ABELIAN))
@pytest.mark.parametrize("J", [0, 1, 2, 3, 4, 5, 8, 11])
def test_abelian_levels_reduce_the_rotation_group_representation(group, J):
    """Irrep counts from the reduction of D^(J) over the rotations
    equivalent to the group's operations (closed-form SO(3) characters)."""
    chars, angles, _ = ABELIAN[group]
    want = [round(sum(c * so3_character(J, a) for c, a in zip(ch, angles))
                  / len(angles)) for ch in chars]
    nl, md, irr, root, vec, ops = functions(group, J)
    assert md == 1 and nl == 2 * J + 1
    assert np.bincount(irr, minlength=len(chars)).tolist() == want


@pytest.mark.parametrize("group", sorted(ABELIAN))
@pytest.mark.parametrize("J", [1, 2, 3, 6, 9])
def test_abelian_levels_carry_the_parity_of_k(group, J):
    """A rotation by pi about z multiplies |J, k> by (-1)^k: each level is
    built from Wang functions of one parity of k, and its irrep's character
    of every operation equivalent to that rotation is (-1)^k."""
    chars, _, rz = ABELIAN[group]
    nl, md, irr, root, vec, ops = functions(group, J)
    k_of_row = np.array([0] + [(i + 1) // 2 for i in range(1, 2 * J + 1)])
    for l in range(nl):
        rows = np.nonzero(np.abs(vec[:, root[l, 0]]) > 1e-12)[0]
        par = set((k_of_row[rows] % 2).tolist())
        assert len(par) == 1, (group, J, l)
        p = par.pop()
        for op in rz:
            assert chars[irr[l]][op] == (-1) ** p, (group, J, l, op)


@pytest.mark.parametrize("group", GROUPS)
@pytest.mark.parametrize("J", [0, 1, 2, 3, 5, 8])
def test_functions_orthonormal_and_blocks_carry_the_irrep(group, J):
    nl, md, irr, root, vec, ops = functions(group, J)
    n = 2 * J + 1
    np.testing.assert_allclose(vec.T @ vec, np.eye(n), atol=1e-13)
    for op in range(NOPER[group]):
        t = ops[op]
        np.testing.assert_allclose(t.T @ t, np.eye(n), atol=1e-12)
        q = vec.T @ t @ vec  # operation in the symmetrised basis
        blocks = [[r for r in root[l] if r >= 0] for l in range(nl)]
        mask = np.zeros((n, n), bool)
        for l, cols in enumerate(blocks):
            sub = q[np.ix_(cols, cols)]
            assert abs(np.trace(sub) - CHARS[group][irr[l]][op]) < 1e-10
            mask[np.ix_(cols, cols)] = True
        assert np.abs(q[~mask]).max(initial=0.0) < 1e-10


def wang_index(b):
    """Wang function (k = 0; then k, tau = 0, 1) of every K-block primitive."""
    k = np.repeat(np.arange(b.J + 1), b.kblock)
    return np.where(k == 0, 0, 2 * k - 1 + (np.asarray(b.tau) & 1))


@pytest.mark.parametrize("jI,jF", [(0, 1), (1, 1), (2, 3), (3, 2), (4, 4)])
def test_no_symmetry_table_is_the_kblock_factor(jI, jF):
    nlI, nlF, md, rot = wigner("C(M)", jI, jF)
    assert (nlI, nlF, md) == (2 * jI + 1, 2 * jF + 1, 1)
    jmax = max(jI, jF)
    tab = orc.threej_table(jmax)
    kt = lambda J: [(0, J % 2)] + [(k, t) for k in range(1, J + 1)
                                   for t in (0, 1)]
    for a, (kI, tI) in enumerate(kt(jI)):
        for b, (kF, tF) in enumerate(kt(jF)):
            v, comp = orc.rot_coeff(jmax, tab, jI, jF, kI, tI, kF, tF)
            want = np.zeros(3)
            if comp >= 0:
                # the oracle carries the kernel's truncated 1/sqrt(2)
                s = 1.0 / 0.70710678118 / math.sqrt(2.0)
                want[comp] = v * (s if (kI != kF and kI * kF != 0) else 1.0)
            np.testing.assert_allclose(rot[0, 0, b, a], want, rtol=0,
                                       atol=1e-14)


@pytest.mark.parametrize("group", GROUPS)
@pytest.mark.parametrize("jI,jF", [(1, 2), (2, 2), (3, 2), (4, 5)])
def test_rotsym_half_ls_is_the_kblock_one_in_the_symmetrised_basis(
        group, jI, jF):
    rng = np.random.default_rng(7 * jI + jF)
    nc = 5
    prob = synth.make_problem([jI, jF] if jI != jF else [jI], nc, "c1")
    bI = prob.bases[0]
    bF = prob.bases[-1]
    jmax = max(jI, jF)
    tab = orc.threej_table(jmax)
    x_kb = rng.standard_normal(bI.nprim)
    h_kb = orc.half_ls(prob, bI, bF, x_kb, tab)
    fI = functions(group, jI)
    fF = functions(group, jF)
    nlI, nlF, md, rot = wigner(group, jI, jF)

    def rotsym_basis(f, b):
        nl, md_, irr, root, vec, _ = f
        levels, degs, vibs, roots = [], [], [], []
        for v in range(nc):
            for l in range(nl):
                for d in range(md_):
                    if root[l, d] >= 0:
                        levels.append(l)
                        degs.append(d)
                        vibs.append(v)
                        roots.append(root[l, d])
        # primitive (v, Wang function) -> position in the K-block basis
        pos = {(int(v), int(w)): i for i, (v, w) in
               enumerate(zip(b.vib, wang_index(b)))}
        return (np.array(levels), np.array(degs), np.array(vibs),
                np.array(roots), vec, pos)

    lvI, dgI, vbI, rtI, EI, posI = rotsym_basis(fI, bI)
    lvF, dgF, vbF, rtF, EF, posF = rotsym_basis(fF, bF)
    nI, nF = 2 * jI + 1, 2 * jF + 1
    # x in the symmetrised basis: x_rs[v, p] = sum_w E_I[w, p] x_kb[v, w]
    x_rs = np.array([sum(EI[w, p] * x_kb[posI[(v, w)]] for w in range(nI))
                     for v, p in zip(vbI, rtI)])
    h_rs = orc.half_ls_rotsym(nlI, nlF, md, dgI, dgF, lvI, lvF, vbI, vbF, 0,
                              nc, prob.mu, nc, rot, x_rs)
    want = np.array([sum(EF[w, q] * h_kb[posF[(v, w)]] for w in range(nF))
                     for v, q in zip(vbF, rtF)])
    scale = np.abs(want).max()
    # the K-block oracle uses the kernel's truncated 1/sqrt(2)
    # (relative difference 9.3e-12)
    np.testing.assert_allclose(h_rs, want, rtol=0, atol=1e-10 * scale)


def test_errors():
    L = _lib.load()
    a, b, m = C.c_int(), C.c_int(), C.c_int()
    assert L.gain_rotsym_wigner(b"C2V(M)", 2, 4, C.byref(a), C.byref(b),
                                C.byref(m), None) == _lib.ERR_ARG
    assert L.gain_rotsym_wigner(b"OH(M)", 2, 2, C.byref(a), C.byref(b),
                                C.byref(m), None) == _lib.ERR_ARG
    assert b"group" in L.gain_last_error()


@pytest.mark.parametrize("group", ["C3V(M)", "TD(M)", "D2H(M)", "C2V(M)"])
def test_rotsym_lines_equal_kblock_lines_in_the_symmetrised_basis(group):
    """The oracle's whole main.cpp:238-338 loop with symmetric-top bases
    (the ROTSYM half line strength, GpuManager.cpp:612-618) on records
    carried from a K-block problem into the symmetrised rotational basis
    gives the K-block problem's transition list and line strengths (up to
    the K-block kernel's truncated 1/sqrt(2))."""
    nc = 4
    kb = synth.make_problem([2, 3], nc, "c1")
    kb.params.erange_lower = (0.0, 1e9)
    rs = synth.make_problem([2, 3], nc, "c1", rotsym=group)
    rs.params = kb.params
    for jind, (bK, bR) in enumerate(zip(kb.bases, rs.bases)):
        nl, md_, irr, root, E, _ = functions(group, bK.J)
        pos = {(int(v), int(w)): i for i, (v, w) in
               enumerate(zip(bK.vib, wang_index(bK)))}
        roots = root[bR.level, bR.deg]
        T = np.zeros((bR.nprim, bK.nprim))
        for p, (v, q) in enumerate(zip(bR.vib, roots)):
            for w in range(2 * bK.J + 1):
                T[p, pos[(int(v), w)]] = E[w, q]
        blk = kb.blocks[(jind, 0)]
        rs.blocks[(jind, 0)] = synth.Block(jind, 0, blk.energy.copy(),
                                           blk.vectors @ T.T, blk.rec_len)
    rows = kb.state_table()
    initials = list(range(0, len(rows), 3))
    trK, _ = orc.run(kb, initials)
    trR, _ = orc.run(rs, initials)
    assert len(trK) > 0
    assert np.array_equal(trK[["iI", "iF", "emitted"]],
                          trR[["iI", "iF", "emitted"]])
    smax = trK["S"].max()
    np.testing.assert_allclose(trR["S"], trK["S"], rtol=1e-9,
                               atol=1e-13 * smax)


@pytest.mark.parametrize("J", [0, 1, 2, 3, 6])
def test_dnh_at_the_reference_order_is_the_trivial_representation(J):
    """DNH(M) at the order GAIN-MPI runs it (sym%N keeps its default 1,
    symmetry.f90:33,746-1006): four operations whose equivalent rotations
    are all the identity, so every rotational function is A1' and the
    functions and tables equal the no-symmetry ones."""
    nl, md, irr, root, vec, ops = functions("DNH(M)", J)
    c = functions("C(M)", J)
    assert (nl, md) == (2 * J + 1, 1) and (irr == 0).all()
    np.testing.assert_array_equal(vec, c[4])
    np.testing.assert_array_equal(ops, np.stack([np.eye(2 * J + 1)] * 4))
    for jF in (J - 1, J, J + 1):
        if jF < 0:
            continue
        np.testing.assert_array_equal(wigner("DNH(M)", J, jF)[3],
                                      wigner("C(M)", J, jF)[3])


def test_dinftyh_is_refused_with_its_reason():
    L = _lib.load()
    a, b, m = C.c_int(), C.c_int(), C.c_int()
    assert L.gain_rotsym_wigner(b"DINFTYH(M)", 2, 2, C.byref(a), C.byref(b),
                                C.byref(m), None) == _lib.ERR_ARG
    assert b"degenerate irreps" in L.gain_last_error()
    assert L.gain_rotsym_wigner(b"C3V(M)", 2, 2, None, None, None,
                                None) == _lib.ERR_ARG
    assert b"null" in L.gain_last_error()
