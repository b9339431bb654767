"""ctypes front end of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The C restatement it loads lives in
oracle.c (each function cites the reference file:line it follows).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class OrcParams(C.Structure):
    _fields_ = [("nrepres", C.c_int), ("maxdeg", C.c_int), ("maxJ", C.c_int),
                ("reduced", C.c_int), ("symmetry_type", C.c_int),
                ("sym_degen", C.c_int * 16), ("isym_do", C.c_int * 16),
                ("isym_pairs", C.c_int * 16), ("igamma_pair", C.c_int * 16),
                ("gns", C.c_double * 16), ("erange_lower", C.c_double * 2),
                ("erange_upper", C.c_double * 2),
                ("freq_window", C.c_double * 2), ("zpe", C.c_double),
                ("temperature", C.c_double), ("partition", C.c_double),
                ("threshold", C.c_double)]


class OrcBasis(C.Structure):
    _fields_ = [("J", C.c_int), ("nprim", C.c_int),
                ("kstart", C.POINTER(C.c_int)), ("kblock", C.POINTER(C.c_int)),
                ("tau", C.POINTER(C.c_int)), ("vib", C.POINTER(C.c_int)),
                ("maxsymcoeffs", C.c_int), ("ijterms", C.POINTER(C.c_int)),
                ("icontr", C.POINTER(C.c_int)), ("matsize", C.c_int),
                ("N", C.POINTER(C.c_int) * 16),
                ("repres", C.POINTER(C.c_double) * 16),
                ("ntot", C.c_int * 16), ("sdeg", C.c_int * 16),
                ("rotsym", C.c_int), ("nlevels", C.c_int), ("rmd", C.c_int),
                ("level", C.POINTER(C.c_int)), ("deg", C.POINTER(C.c_int)),
                ("wig", C.POINTER(C.c_double) * 3)]


class OrcStates(C.Structure):
    _fields_ = [("n", C.c_int), ("jind", C.POINTER(C.c_int)),
                ("jval", C.POINTER(C.c_int)), ("igamma", C.POINTER(C.c_int)),
                ("ndeg", C.POINTER(C.c_int)), ("ilevel", C.POINTER(C.c_int)),
                ("rec_len", C.POINTER(C.c_int)),
                ("energy", C.POINTER(C.c_double)),
                ("vectors", C.POINTER(C.c_double)),
                ("offset", C.POINTER(C.c_int64))]


class OrcTransition(C.Structure):
    _fields_ = [("iI", C.c_int), ("iF", C.c_int), ("emitted", C.c_int),
                ("S", C.c_double), ("A", C.c_double), ("I", C.c_double),
                ("ls", C.c_double * 9)]


TRANSITION_DTYPE = np.dtype([("iI", np.int32), ("iF", np.int32),
                             ("emitted", np.int32), ("S", np.float64),
                             ("A", np.float64), ("I", np.float64),
                             ("ls", np.float64, (9,))], align=True)


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_three_j.restype = C.c_double
        L.orc_three_j.argtypes = [C.c_int] * 6
        L.orc_threej_index.restype = C.c_int
        L.orc_threej_index.argtypes = [C.c_int] * 5
        L.orc_threej_table.argtypes = [C.c_int, _f64p]
        L.orc_rot_coeff.restype = C.c_double
        L.orc_rot_coeff.argtypes = [C.c_int, _f64p] + [C.c_int] * 6 + [
            C.POINTER(C.c_int)]
        L.orc_expand.argtypes = [C.c_int] * 6 + [_i32p, _i32p, _i32p, _f64p,
                                                 _f64p, _f64p]
        L.orc_half_ls_kblock.argtypes = (
            [C.c_int] * 5 + [_i32p] * 4 + [C.c_int] * 5 + [_i32p, _f64p,
                                                         C.c_int, _f64p,
                                                         _f64p, C.c_int,
                                                         _f64p])
        L.orc_half_ls.argtypes = (
            [C.c_int, C.c_int] + [_i32p] * 4 + [C.c_int, C.c_int] +
            [_i32p] * 4 + [_f64p, C.c_int, C.c_int, C.c_int, _f64p, C.c_int,
                           _f64p, _f64p])
        L.orc_half_ls_rotsym.argtypes = (
            [C.c_int] * 5 + [_i32p] * 6 + [C.c_int] * 3 +
            [_f64p, C.c_int, _f64p, _f64p, _f64p])
        L.orc_set_plain_fp64.restype = None
        L.orc_set_plain_fp64.argtypes = [C.c_int]
        L.orc_dot.restype = C.c_double
        L.orc_dot.argtypes = [C.c_int64, _f64p, _f64p]
        L.orc_gamma_pairs.restype = C.c_int
        L.orc_gamma_pairs.argtypes = [C.POINTER(OrcParams)]
        for f in ("orc_filter_lower", "orc_filter_upper"):
            getattr(L, f).restype = C.c_int
            getattr(L, f).argtypes = [C.POINTER(OrcParams), C.c_double,
                                      C.c_int]
        L.orc_filter_intensity.restype = C.c_int
        L.orc_filter_intensity.argtypes = [C.POINTER(OrcParams), C.c_int,
                                           C.c_int, C.c_double, C.c_int,
                                           C.c_int, C.c_double]
        L.orc_degeneracy_filter.restype = C.c_int
        L.orc_degeneracy_filter.argtypes = [C.POINTER(OrcParams)] + [
            C.c_int] * 4
        L.orc_line.restype = C.c_int
        L.orc_line.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_int,
                               C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                               C.c_double, _f64p] + [C.POINTER(C.c_double)] * 3
        L.orc_format_line.restype = C.c_int
        L.orc_format_line.argtypes = [C.POINTER(OrcParams), C.c_char_p,
                                      C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int, C.c_double, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_double, _f64p,
                                      C.c_int, C.c_int]
        L.orc_run.restype = C.c_int64
        L.orc_run.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcBasis),
                              C.c_int, C.POINTER(OrcStates), _f64p, C.c_int,
                              C.c_int, _i32p, C.c_int, C.c_void_p, C.c_int64,
                              C.c_int]
        _lib = L
    return _lib


# ---------------------------------------------------------------- helpers
def params_struct(p) -> OrcParams:
    s = OrcParams()
    s.nrepres = p.nrepres
    s.maxdeg = p.maxdeg
    s.maxJ = p.maxJ
    s.reduced = int(bool(p.reduced))
    s.symmetry_type = p.symmetry_type
    for g in range(p.nrepres):
        s.sym_degen[g] = p.sym_degen[g]
        s.isym_do[g] = int(p.isym_do[g])
        s.isym_pairs[g] = p.isym_pairs[g]
        s.igamma_pair[g] = p.igamma_pair[g]
        s.gns[g] = p.gns[g]
    s.erange_lower[:] = list(p.erange_lower)
    s.erange_upper[:] = list(p.erange_upper)
    s.freq_window[:] = list(p.freq_window)
    s.zpe, s.temperature = p.zpe, p.temperature
    s.partition, s.threshold = p.partition, p.threshold
    return s


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def three_j(j1, j2, j3, m1, m2, m3) -> float:
    return lib().orc_three_j(j1, j2, j3, m1, m2, m3)


def threej_table(jmax: int) -> np.ndarray:
    t = np.zeros((jmax + 1) * (jmax + 1) * 9)
    lib().orc_threej_table(jmax, t)
    return t


def threej_index(jmax, jI, kI, jF, kF) -> int:
    return lib().orc_threej_index(jmax, jI, kI, jF, kF)


def rot_coeff(jmax, tab, jI, jF, kI, tauI, kF, tauF):
    a = C.c_int(0)
    v = lib().orc_rot_coeff(jmax, tab, jI, jF, kI, tauI, kF, tauF,
                            C.byref(a))
    return v, a.value


def expand(basis, gamma, ideg, vin) -> np.ndarray:
    out = np.zeros(basis.nprim)
    lib().orc_expand(basis.nprim, gamma, basis.maxsymcoeffs, ideg,
                     basis.sdeg[gamma], basis.ntot[gamma], basis.ijterms,
                     basis.icontr, basis.N[gamma],
                     np.ascontiguousarray(basis.repres[gamma], np.float64),
                     np.ascontiguousarray(vin, np.float64), out)
    return out


def half_ls(prob, bI, bF, x, tab=None, startF=0, endF=None) -> np.ndarray:
    """Half line strength J_I -> J_F with one dipole piece [startF,endF)."""
    jmax = prob.jmax
    if tab is None:
        tab = threej_table(jmax)
    nc = prob.nc
    endF = nc if endF is None else endF
    mu = prob.mu
    if startF != 0 or endF != nc:
        mu = np.ascontiguousarray(prob.mu[:, :, startF:endF])
    h = np.zeros(bF.nprim)
    lib().orc_half_ls(bI.J, bI.nprim, bI.kstart, bI.kblock, bI.tau, bI.vib,
                      bF.J, bF.nprim, bF.kstart, bF.kblock, bF.tau, bF.vib,
                      np.ascontiguousarray(mu).reshape(-1), nc, startF, endF,
                      tab, jmax, np.ascontiguousarray(x, np.float64), h)
    return h


def half_ls_rotsym(nlevelI, nlevelF, maxdeg, degI, degF, levelI, levelF,
                   vibI, vibF, startF, endF, dipole, dip_stride, wigner, x,
                   out=None) -> np.ndarray:
    """Symmetric-top half line strength (dipole_GPU.cu:641-723).

    dipole: the piece [3, dip_stride, endF-startF] (final index fastest);
    wigner: [maxdeg, maxdeg, nlevelF, nlevelI, 3] (last index fastest).
    Entries of `out` whose final vibrational index is outside
    [startF, endF) keep their value."""
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    degI, degF, levelI, levelF, vibI, vibF = map(
        i32, (degI, degF, levelI, levelF, vibI, vibF))
    h = np.zeros(vibF.shape[0]) if out is None else np.array(out, np.float64)
    lib().orc_half_ls_rotsym(
        vibF.shape[0], vibI.shape[0], nlevelI, nlevelF, maxdeg, degI, degF,
        levelI, levelF, vibI, vibF, startF, endF, endF - startF,
        np.ascontiguousarray(dipole, np.float64).reshape(-1), dip_stride,
        np.ascontiguousarray(wigner, np.float64).reshape(-1),
        np.ascontiguousarray(x, np.float64), h)
    return h


def dot(x, y) -> float:
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().orc_dot(x.shape[0], x, y)


def basis_struct(b, keep) -> OrcBasis:
    s = OrcBasis()
    s.J, s.nprim = b.J, b.nprim
    for name in ("kstart", "kblock", "tau", "vib", "ijterms", "icontr"):
        arr = np.ascontiguousarray(getattr(b, name), np.int32)
        keep.append(arr)
        setattr(s, name, _ip(arr))
    s.maxsymcoeffs, s.matsize = b.maxsymcoeffs, b.matsize
    for g in range(len(b.N)):
        n = np.ascontiguousarray(b.N[g], np.int32)
        r = np.ascontiguousarray(b.repres[g], np.float64)
        if r.size == 0:
            r = np.zeros(1)
        keep += [n, r]
        s.N[g] = _ip(n)
        s.repres[g] = _dp(r)
        s.ntot[g] = b.ntot[g]
        s.sdeg[g] = b.sdeg[g]
    if getattr(b, "rotsym", False):
        s.rotsym, s.nlevels, s.rmd = 1, b.nlevels, b.rmd
        for name in ("level", "deg"):
            arr = np.ascontiguousarray(getattr(b, name), np.int32)
            keep.append(arr)
            setattr(s, name, _ip(arr))
        for dj, rot in (b.wig or {}).items():
            arr = np.ascontiguousarray(rot, np.float64)
            keep.append(arr)
            s.wig[dj] = _dp(arr)
    return s


def run(prob, initial_states, nthreads: int = 0, finals=None,
        plain_fp64: bool = False, stats=None):
    """main.cpp:238-338 over `initial_states` (indices into the energy-sorted
    state table).  `finals` (optional) restricts the state table to those
    indices plus the initials (used to bound CPU time on large problems).
    `plain_fp64` switches the accumulations from double-double to plain FP64
    (CPU timing only); `stats` (a dict) receives the seconds spent in the
    compute call, without the record staging.
    Returns (structured transitions array, state table rows)."""
    rows = prob.state_table()
    keep_idx = None
    if finals is not None:
        keep_idx = sorted(set(finals) | set(initial_states))
        remap = {old: new for new, old in enumerate(keep_idx)}
        rows_k = [rows[i] for i in keep_idx]
        init_k = [remap[i] for i in initial_states]
    else:
        rows_k, init_k = rows, list(initial_states)
    n = len(rows_k)
    jind = np.array([r[1] for r in rows_k], np.int32)
    igam = np.array([r[2] for r in rows_k], np.int32)
    lvl = np.array([r[3] for r in rows_k], np.int32)
    jval = np.array([prob.bases[j].J for j in jind], np.int32)
    ndeg = np.array([prob.params.sym_degen[g] for g in igam], np.int32)
    energy = np.array([r[0] for r in rows_k], np.float64)
    rec = np.array([prob.blocks[(j, g)].rec_len for j, g in zip(jind, igam)],
                   np.int32)
    offset = np.concatenate([[0], np.cumsum(rec.astype(np.int64))[:-1]])
    vectors = np.empty(int(rec.astype(np.int64).sum()) + 1)
    for i, (j, g, s) in enumerate(zip(jind, igam, lvl)):
        vectors[offset[i]:offset[i] + rec[i]] = prob.blocks[(j, g)].vector(s)
    offset = offset.astype(np.int64)
    st = OrcStates()
    st.n = n
    keep = [jind, igam, lvl, jval, ndeg, energy, rec, offset, vectors]
    st.jind, st.jval, st.igamma = _ip(jind), _ip(jval), _ip(igam)
    st.ndeg, st.ilevel, st.rec_len = _ip(ndeg), _ip(lvl), _ip(rec)
    st.energy, st.vectors = _dp(energy), _dp(vectors)
    st.offset = offset.ctypes.data_as(C.POINTER(C.c_int64))
    bases = (OrcBasis * len(prob.bases))()
    for i, b in enumerate(prob.bases):
        bases[i] = basis_struct(b, keep)
    ps = params_struct(prob.params)
    init = np.array(init_k, np.int32)
    cap = max(1, len(init_k) * n)
    out = np.zeros(cap, TRANSITION_DTYPE)
    assert out.dtype.itemsize == C.sizeof(OrcTransition)
    mu = np.ascontiguousarray(prob.mu).reshape(-1)
    lib().orc_set_plain_fp64(1 if plain_fp64 else 0)
    t0 = time.perf_counter()
    try:
        cnt = lib().orc_run(C.byref(ps), bases, len(prob.bases), C.byref(st),
                            mu, prob.nc, prob.jmax, init, len(init_k),
                            out.ctypes.data_as(C.c_void_p), cap, nthreads)
    finally:
        lib().orc_set_plain_fp64(0)
    if stats is not None:
        stats["seconds"] = time.perf_counter() - t0
    if cnt < 0:
        raise RuntimeError("oracle run failed")
    out = out[:cnt].copy()
    if keep_idx is not None:
        ki = np.array(keep_idx, np.int32)
        out["iI"] = ki[out["iI"]]
        out["iF"] = ki[out["iF"]]
    return out, rows
