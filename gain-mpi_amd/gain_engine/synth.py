"""Seeded synthetic line-strength problems (host side, numpy).

The reference reads TROVE checkpoint files; none ship with it, so every
workload here is generated.  The generated objects carry exactly the
information the reference's host classes hand to its device layer:

* a per-J primitive basis: K-blocks, tau and vibrational (J=0 contracted)
  index per primitive, K-major ordering (TroveBasisSet.cpp:454-495,658-681);
* per-J inflation data: ijterms / icontr / N / repres per irrep
  (TroveBasisSet.cpp:688-709, GpuManager.cpp:381-422);
* the dipole mu in the device piece layout mu[alpha][v_I][v_F]
  (TroveDipole.cpp:110, one piece covering every v_F);
* a state table (energy-sorted, States.h:6-25) and eigenvector records, one
  record of rec_len doubles per state (EigenVector.cpp:152-178).

Symmetry schemes (groups of consecutive primitives mixed by a fixed
orthogonal matrix whose rows are split among the irreps): "c1"; "c2v4"
(Hadamard quartets, 4 one-dimensional irreps); "pair2" (symmetric and
antisymmetric pairs); "c3v3" (triplets -> A1, A2 empty, E; C3v selection and
reduced degeneracy rules); "td5" (decades -> A1, A2, E, F1, F2 with the
degeneracies, pairs A1-A2, E-E, F1-F2 and reduced degeneracy rules of Td,
States.cpp:108-115).

Distributions: mu_alpha entries N(0, 0.1) damped by exp(-|a-b|/50), z
symmetric, x and y antisymmetric; eigenvectors orthonormal (QR of a
Gaussian, or a Kronecker product of two such factors for blocks too large to
factorise); energies uniform and sorted.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, List, Optional, Tuple

import numpy as np

SEED_MU, SEED_CI, SEED_CF, SEED_E = 1, 2, 3, 4


@dataclasses.dataclass
class Basis:
    """One J: primitive metadata plus inflation data for every irrep."""

    J: int
    nprim: int
    kstart: np.ndarray  # int32[J+1]
    kblock: np.ndarray  # int32[J+1]
    tau: np.ndarray  # int32[nprim]
    vib: np.ndarray  # int32[nprim]
    maxsymcoeffs: int
    ijterms: np.ndarray  # int32[maxsymcoeffs * nrepres]
    icontr: np.ndarray  # int32[2 * nprim]: case, then lambda
    matsize: int
    N: List[np.ndarray]  # per irrep int32[maxsymcoeffs]
    repres: List[np.ndarray]  # per irrep float64[ntot*sdeg*matsize]
    ntot: List[int]
    sdeg: List[int]
    # symmetric-top (ROTSYM) basis: rotational level and degenerate
    # component per primitive, and the rotational tables to J_F = J + dj - 1
    # (dj = 0, 1, 2) present in the J list
    rotsym: bool = False
    nlevels: int = 0
    rmd: int = 0
    level: Optional[np.ndarray] = None
    deg: Optional[np.ndarray] = None
    wig: Optional[Dict[int, np.ndarray]] = None

    def kval(self) -> np.ndarray:
        if self.rotsym:
            return np.zeros(self.nprim, np.int32)
        k = np.empty(self.nprim, np.int32)
        for K in range(self.J + 1):
            k[self.kstart[K]:self.kstart[K] + self.kblock[K]] = K
        return k


@dataclasses.dataclass
class Params:
    """Input.h getters used by States / Output."""

    nrepres: int
    sym_degen: List[int]
    gns: List[float]
    isym_pairs: List[int]
    maxJ: int
    symmetry_type: int = 1
    reduced: bool = False
    isym_do: Optional[List[int]] = None
    erange_lower: Tuple[float, float] = (0.0, 10000.0)
    erange_upper: Tuple[float, float] = (0.0, 25000.0)
    freq_window: Tuple[float, float] = (0.0, 15000.0)
    zpe: float = 0.0
    temperature: float = 296.0
    partition: float = 1000.0
    threshold: float = 0.0
    igamma_pair: Optional[List[int]] = None

    @property
    def maxdeg(self) -> int:
        return max(self.sym_degen)

    def __post_init__(self):
        if self.isym_do is None:
            self.isym_do = [1] * self.nrepres
        if self.igamma_pair is None:
            self.igamma_pair = gamma_pairs(self.isym_pairs, self.gns)


def gamma_pairs(isym_pairs, gns) -> List[int]:
    """Pair of each irrep under the selection rules (TroveInput.cpp:202-236)."""
    n = len(isym_pairs)
    out = []
    for gI in range(n):
        pair, ng = gI, 0
        for gF in range(n):
            if gI != gF and isym_pairs[gI] == isym_pairs[gF]:
                pair = gF
                ng += 1
                if ng > 1:
                    raise ValueError("selection rules do not come in pairs")
        if gns[gI] != gns[pair]:
            raise ValueError("selection rules do not agree with gns")
        out.append(pair)
    return out


@dataclasses.dataclass
class Block:
    """All states of one (J index, irrep): energies and eigenvector records."""

    jind: int
    gamma: int
    energy: np.ndarray  # float64[n], sorted ascending
    vectors: Optional[np.ndarray]  # float64[n, rec_len] or None (device-made)
    rec_len: int
    kron: Optional[Tuple[np.ndarray, np.ndarray]] = None  # (Qa, Qb) factors

    @property
    def n(self) -> int:
        return int(self.energy.shape[0])

    def vector(self, s: int) -> np.ndarray:
        if self.vectors is not None:
            return self.vectors[s]
        qa, qb = self.kron
        b = qb.shape[0]
        k, l = divmod(s, b)
        return np.outer(qa[:, k], qb[:, l]).reshape(-1)

    def rows(self, lo: int, hi: int) -> np.ndarray:
        if self.vectors is not None:
            return self.vectors[lo:hi]
        return np.stack([self.vector(s) for s in range(lo, hi)])


@dataclasses.dataclass
class Problem:
    nc: int
    mu: np.ndarray  # float64[3, nc(v_I), nc(v_F)]
    bases: List[Basis]
    params: Params
    blocks: Dict[Tuple[int, int], Block]

    @property
    def jvals(self) -> List[int]:
        return [b.J for b in self.bases]

    @property
    def jmax(self) -> int:
        return max(self.jvals)

    def state_table(self):
        """Energy-sorted global state table (TroveStates.cpp:162)."""
        rows = []
        for (jind, g), blk in sorted(self.blocks.items()):
            for s in range(blk.n):
                rows.append((float(blk.energy[s]), jind, g, s))
        rows.sort(key=lambda r: r[0])
        return rows


# --------------------------------------------------------------------------
def make_mu(nc: int, seed: int = SEED_MU) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = np.arange(nc)
    damp = np.exp(-np.abs(a[:, None] - a[None, :]) / 50.0)
    mu = np.empty((3, nc, nc))
    for s in range(3):
        m = rng.standard_normal((nc, nc)) * 0.1 * damp
        up = np.triu(m, 1)
        if s == 2:  # z symmetric
            mat = up + up.T + np.diag(np.diag(m))
        else:  # x, y antisymmetric, zero diagonal
            mat = up - up.T
        # mat[vF, vI]; device layout is mu[alpha][vI][vF]
        mu[s] = mat.T
    return np.ascontiguousarray(mu)


def _scheme(name: str):
    """Symmetry scheme: (nrepres, sdeg, group size, per-irrep component rows)."""
    if name == "c1":
        return 1, [1], 1, [np.ones((1, 1))]
    if name == "c2v4":  # Hadamard quartets, 4 one-dimensional irreps
        h = np.array([[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, 1, -1],
                      [1, -1, -1, 1]], float) / 2.0
        return 4, [1, 1, 1, 1], 4, [h[g:g + 1] for g in range(4)]
    if name == "pair2":  # symmetric / antisymmetric pairs
        r = np.array([[1, 1], [1, -1]], float) / np.sqrt(2.0)
        return 2, [1, 1], 2, [r[g:g + 1] for g in range(2)]
    if name == "c3v3":  # triplets -> A1 (1), A2 (empty), E (2 components)
        th = np.linspace(0, 2 * np.pi, 4)[:3]
        o = np.stack([np.ones(3) / np.sqrt(3.0),
                      np.sqrt(2.0 / 3.0) * np.cos(th),
                      np.sqrt(2.0 / 3.0) * np.sin(th)])
        return 3, [1, 1, 2], 3, [o[0:1], np.zeros((1, 3)), o[1:3]]
    if name == "td5":  # decades -> A1, A2, E (2), F1 (3), F2 (3) (Td order)
        q = orthonormal(10, np.random.default_rng(5050))
        return 5, [1, 1, 2, 3, 3], 10, [q[0:1], q[1:2], q[2:4], q[4:7],
                                        q[7:10]]
    raise ValueError(name)


def make_basis(J: int, nc: int, scheme: str = "c1") -> Basis:
    """K-major (K=0..J), then tau, then v; K=0 carries tau = J mod 2."""
    nrep, sdeg, gsize, rows = _scheme(scheme)
    if nc % gsize:
        raise ValueError("nc must be a multiple of the symmetry group size")
    kstart, kblock, tau, vib = [], [], [], []
    pos = 0
    for K in range(J + 1):
        taus = [J % 2] if K == 0 else [0, 1]
        kstart.append(pos)
        kblock.append(len(taus) * nc)
        for t in taus:
            tau.extend([t] * nc)
            vib.extend(range(nc))
        pos += len(taus) * nc
    return _with_inflation(J, pos, np.array(kstart, np.int32),
                           np.array(kblock, np.int32),
                           np.array(tau, np.int32), np.array(vib, np.int32),
                           scheme)


def rotsym_levels(group: str, J: int):
    """(nlevels, maxdeg, root_of[nlevels, maxdeg]) of the symmetrised
    rotational functions of one J (the library's generator,
    gain_rotsym_functions); root_of is -1 past a level's degeneracy."""
    import ctypes as C

    from . import _lib
    L = _lib.load()
    nl, md = C.c_int(), C.c_int()
    _lib.check(L.gain_rotsym_functions(group.encode(), J, C.byref(nl),
                                       C.byref(md), None, None, None, None))
    root = np.zeros(nl.value * md.value, np.int32)
    _lib.check(L.gain_rotsym_functions(group.encode(), J, C.byref(nl),
                                       C.byref(md), None, root.ctypes.data,
                                       None, None))
    return nl.value, md.value, root.reshape(nl.value, md.value)


def rotsym_wigner(group: str, jI: int, jF: int) -> np.ndarray:
    """The rotational table of a J_I -> J_F pair (gain_rotsym_wigner):
    3 x nlevelsI x nlevelsF x maxdeg x maxdeg, first index fastest."""
    import ctypes as C

    from . import _lib
    L = _lib.load()
    n = [C.c_int(), C.c_int(), C.c_int()]
    _lib.check(L.gain_rotsym_wigner(group.encode(), jI, jF,
                                    *[C.byref(x) for x in n], None))
    rot = np.zeros(3 * n[0].value * n[1].value * n[2].value ** 2)
    _lib.check(L.gain_rotsym_wigner(group.encode(), jI, jF,
                                    *[C.byref(x) for x in n],
                                    rot.ctypes.data))
    return rot


def make_basis_rotsym(J: int, nc: int, group: str,
                      scheme: str = "c1") -> Basis:
    """Symmetric-top basis of one J (the do_rotsym basis of
    TroveBasisSet.cpp:480-483): primitives ordered vibrational index
    slowest, then rotational level, then degenerate component; the
    inflation scheme mixes consecutive primitives as in make_basis."""
    nl, md, root = rotsym_levels(group, J)
    lv, dg = np.nonzero(root >= 0)
    nprim = nc * len(lv)
    _, _, gsize, _ = _scheme(scheme)
    if nprim % gsize:
        raise ValueError("primitive count must be a multiple of the "
                         "symmetry group size")
    b = _with_inflation(J, nprim, np.zeros(J + 1, np.int32),
                        np.zeros(J + 1, np.int32),
                        np.zeros(nprim, np.int32),
                        np.repeat(np.arange(nc, dtype=np.int32), len(lv)),
                        scheme)
    b.rotsym, b.nlevels, b.rmd = True, nl, md
    b.level = np.tile(lv.astype(np.int32), nc)
    b.deg = np.tile(dg.astype(np.int32), nc)
    return b


def _with_inflation(J, nprim, kstart, kblock, tau, vib, scheme) -> Basis:
    """Basis with the inflation data of `scheme` over groups of consecutive
    primitives."""
    nrep, sdeg, gsize, rows = _scheme(scheme)
    ncase = nprim // gsize
    case = np.arange(nprim) // gsize
    lam = np.arange(nprim) % gsize
    icontr = np.concatenate([case, lam]).astype(np.int32)
    N, repres, ntot = [], [], []
    ijterms = np.zeros(ncase * nrep, np.int32)
    for g in range(nrep):
        has = 0 if not np.any(rows[g]) else 1
        n_g = np.full(ncase, has, np.int32)
        ijterms[g * ncase:(g + 1) * ncase] = np.concatenate(
            [[0], np.cumsum(n_g)[:-1]]).astype(np.int32)
        nt = int(n_g.sum())
        # repres[t + d*Ntot + lambda*sdeg*Ntot]
        rep = np.zeros((gsize, sdeg[g], max(nt, 1)))
        for d in range(sdeg[g]):
            for lm in range(gsize):
                rep[lm, d, :nt] = rows[g][d, lm] if has else 0.0
        N.append(n_g)
        repres.append(np.ascontiguousarray(rep[:, :, :nt]).reshape(-1)
                      if nt else np.zeros(0))
        ntot.append(nt)
    return Basis(J=J, nprim=nprim, kstart=kstart, kblock=kblock, tau=tau,
                 vib=vib, maxsymcoeffs=ncase, ijterms=ijterms, icontr=icontr,
                 matsize=gsize, N=N, repres=repres, ntot=ntot, sdeg=sdeg)


def orthonormal(n: int, rng: np.random.Generator) -> np.ndarray:
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))[None, :]


def kron_factors(n: int) -> Tuple[int, int]:
    """Split n = a*b with a, b as close as possible (a >= b)."""
    b = int(np.floor(np.sqrt(n)))
    while n % b:
        b -= 1
    return n // b, b


def make_block(jind: int, gamma: int, ntot: int, rng_e, rng_c,
               emax: float = 10000.0, kron: bool = False,
               nstates: Optional[int] = None) -> Block:
    n = ntot if nstates is None else nstates
    energy = np.sort(rng_e.uniform(0.0, emax, n))
    if ntot == 0:
        return Block(jind, gamma, energy[:0], np.zeros((0, 0)), 0)
    if kron:
        a, b = kron_factors(ntot)
        qa, qb = orthonormal(a, rng_c), orthonormal(b, rng_c)
        return Block(jind, gamma, energy, None, ntot, kron=(qa, qb))
    q = orthonormal(ntot, rng_c)
    vec = np.ascontiguousarray(q.T[:n])  # record s = column s of Q
    return Block(jind, gamma, energy, vec, ntot)


def make_problem(jvals, nc: int, scheme: str = "c1", blocks=None,
                 kron: bool = False, emax: float = 10000.0,
                 params: Optional[Params] = None, seed_offset: int = 0,
                 nstates: Optional[Dict[Tuple[int, int], int]] = None,
                 rotsym: Optional[str] = None) -> Problem:
    """Problem over a J list.  `blocks` selects which (J index, irrep) blocks
    carry states (default: every non-empty one).  `rotsym` names a
    molecular symmetry group: symmetric-top bases with that group's
    rotational levels and tables instead of K-block bases."""
    if rotsym is None:
        bases = [make_basis(J, nc, scheme) for J in jvals]
    else:
        bases = [make_basis_rotsym(J, nc, rotsym, scheme) for J in jvals]
        for b in bases:
            b.wig = {jF - b.J + 1: rotsym_wigner(rotsym, b.J, jF)
                     for jF in jvals if abs(jF - b.J) <= 1}
    nrep, sdeg, _, _ = _scheme(scheme)
    if params is None:
        pairs = {"c1": [1], "c2v4": [1, 1, 2, 2], "pair2": [1, 1],
                 "c3v3": [1, 1, 2], "td5": [1, 1, 2, 3, 3]}[scheme]
        stype = {"c3v3": 2, "td5": 3}.get(scheme, 1)
        params = Params(nrepres=nrep, sym_degen=list(sdeg),
                        gns=[1.0] * nrep, isym_pairs=pairs, maxJ=max(jvals),
                        symmetry_type=stype, reduced=stype > 1)
    mu = make_mu(nc, SEED_MU + seed_offset)
    rng_e = np.random.default_rng(SEED_E + seed_offset)
    rng_c = np.random.default_rng(SEED_CF + seed_offset)
    out = {}
    for jind, b in enumerate(bases):
        for g in range(nrep):
            if blocks is not None and (jind, g) not in blocks:
                continue
            if b.ntot[g] == 0:
                continue
            ns = None if nstates is None else nstates.get((jind, g))
            out[(jind, g)] = make_block(jind, g, b.ntot[g], rng_e, rng_c,
                                        emax, kron, ns)
    return Problem(nc=nc, mu=mu, bases=bases, params=params, blocks=out)
