"""Shared test helpers: irregular bases, textbook rotational factor and a
dense brute-force half line strength."""
import math

import numpy as np

from gain_engine import synth

# Reciprocal of the reference's truncated 1/sqrt(2) literal (defines.h:23):
# dividing by it reproduces the kernel's multiplication by 0.70710678118.
SQRT2_REF_INV = 1.0 / 0.70710678118


def textbook_rot(jI, jF, kI, tauI, kF, tauF, f3j, sqrt2=math.sqrt(2.0)):
    """Rotational factor in the branchy textbook form of
    trove_wrappers.f90:329-381 (components 1=x, 2=y, 3=z -> 0, 1, 2).
    Returns (alpha, value) or (None, 0.0) when |dK| > 1."""
    if abs(kI - kF) > 1:
        return None, 0.0
    sigmaI = (kI % 3) * tauI
    f_I = (-1.0) ** (sigmaI + kI)
    sigmaF = (kF % 3) * tauF
    f_IF = f_I * (-1.0) ** sigmaF
    if kF == kI:
        return 2, (tauF - tauI) * f3j * f_IF
    if tauF != tauI:
        v = (kF - kI) * (tauF - tauI) * f3j * f_IF
        if kI * kF != 0:
            v = v / sqrt2
        return 0, v
    v = -f3j * f_IF
    if kI * kF != 0:
        v = v / sqrt2
    return 1, v


def irregular_basis(J, nc, rng, max_per_block=None):
    """A C1 basis whose K blocks have random sizes, random tau per primitive
    and random (possibly repeated) vibrational indices."""
    kstart, kblock, tau, vib = [], [], [], []
    pos = 0
    mx = max_per_block or 2 * nc
    for K in range(J + 1):
        n = int(rng.integers(0, mx + 1))
        kstart.append(pos)
        kblock.append(n)
        tau.extend(rng.integers(0, 2, n).tolist())
        vib.extend(rng.integers(0, nc, n).tolist())
        pos += n
    nprim = pos
    b = synth.make_basis(0, 1, "c1")  # template for the C1 inflation fields
    return synth.Basis(
        J=J, nprim=nprim, kstart=np.array(kstart, np.int32),
        kblock=np.array(kblock, np.int32), tau=np.array(tau, np.int32),
        vib=np.array(vib, np.int32), maxsymcoeffs=max(nprim, 1),
        ijterms=np.arange(max(nprim, 1), dtype=np.int32),
        icontr=np.concatenate([np.arange(nprim), np.zeros(nprim)]).astype(
            np.int32),
        matsize=1, N=[np.ones(max(nprim, 1), np.int32)],
        repres=[np.ones(nprim)], ntot=[nprim], sdeg=[1])


def dense_half_ls(bI, bF, mu, x, tab, jmax, threej_index):
    """h = D x with D assembled element by element from the textbook form."""
    kI_all, kF_all = bI.kval(), bF.kval()
    h = np.zeros(bF.nprim)
    for rp in range(bF.nprim):
        kF, tF, vF = int(kF_all[rp]), int(bF.tau[rp]) & 1, int(bF.vib[rp])
        terms = []
        for r in range(bI.nprim):
            kI, tI, vI = int(kI_all[r]), int(bI.tau[r]), int(bI.vib[r])
            if abs(kI - kF) > 1:
                continue
            f3j = tab[threej_index(jmax, bI.J, kI, bF.J, kF)]
            a, c = textbook_rot(bI.J, bF.J, kI, tI, kF, tF, f3j,
                                sqrt2=SQRT2_REF_INV)
            terms.append(c * mu[a, vI, vF] * x[r])
        h[rp] = math.fsum(terms)
    return h


def problem_with_bases(bases, nc, seed=0, nrep_blocks=None):
    """C1 problem over explicit (possibly irregular) bases."""
    prob = synth.make_problem([b.J for b in bases], nc, "c1",
                              seed_offset=seed)
    prob.bases = list(bases)
    rng_e = np.random.default_rng(100 + seed)
    rng_c = np.random.default_rng(200 + seed)
    prob.blocks = {}
    for jind, b in enumerate(bases):
        if b.ntot[0] > 0:
            prob.blocks[(jind, 0)] = synth.make_block(jind, 0, b.ntot[0],
                                                      rng_e, rng_c)
    prob.params.maxJ = max(b.J for b in bases)
    return prob


def oracle_map(tr):
    return {(int(t["iI"]), int(t["iF"])): t for t in tr}


def engine_lines(eng, prob, initials, mode="plain", precision=0,
                 batch=None):
    """Run the engine for every initial state (global state-table indices)
    against every registered block; returns {(I, F): record} for pairs whose
    pass flag is set, with the same fields as the oracle transitions."""
    rows = prob.state_table()
    gidx = {(r[1], r[2], r[3]): i for i, r in enumerate(rows)}
    out = {}
    by_block = {}
    for I in initials:
        by_block.setdefault((rows[I][1], rows[I][2]), []).append(I)
    for (jI, gI), Is in by_block.items():
        for (jF, gF), blk in prob.blocks.items():
            if mode == "plain":
                res = []
                for I in Is:
                    r = eng.plain_lines(jI, gI, rows[I][3], jF, gF)
                    res.append({k: v[None] for k, v in r.items()})
                res = {k: np.concatenate([r[k] for r in res]) for k in res[0]}
            else:
                levels = [rows[I][3] for I in Is]
                res = eng.batched_lines(jI, gI, levels, jF, gF, precision)
            for i, I in enumerate(Is):
                for f in np.nonzero(res["flags"][i] & 1)[0]:
                    F = gidx[(jF, gF, int(f))]
                    out[(I, F)] = {"S": res["S"][i, f], "A": res["A"][i, f],
                                   "I": res["I"][i, f],
                                   "ls": res["ls"][i, f],
                                   "emitted": int(res["flags"][i, f] >> 1)}
    return out


def compare_lines(got, ref, rel=1e-12, what=("S", "A", "I")):
    """Transition lists identical; values within `rel` on strong lines and
    within rel * max on weak ones.  Returns the max relative error on
    strong lines."""
    assert set(got) == set(ref), (len(set(got) - set(ref)),
                                  len(set(ref) - set(got)))
    worst = 0.0
    if not ref:
        return worst
    for key in what:
        r = np.array([ref[k][key] for k in ref])
        g = np.array([got[k][key] for k in ref])
        big = np.abs(r).max()
        strong = np.abs(r) > 1e-6 * big
        err = np.abs(g - r)
        if strong.any():
            e = (err[strong] / np.abs(r[strong])).max()
            worst = max(worst, e)
            assert e <= rel, (key, e)
        assert (err <= rel * big + 1e-300).all(), key
    em_r = {k for k in ref if ref[k]["emitted"]}
    em_g = {k for k in got if got[k]["emitted"]}
    assert em_r == em_g
    return worst
