"""Readers of the TROVE checkpoint files (host side, through the C ABI):
j0_extfield.chk, j0eigen_quanta<J>.chk, j0eigen_descr<J>_<g>.chk and
j0eigen_vectors<J>_<g>.chk, in the formats the reference reads them
(TroveDipole.cpp:9-119, TroveBasisSet.cpp:27-709, TroveStates.cpp:15-131,
EigenVector.cpp:36-88)."""
from __future__ import annotations

import ctypes as C
from typing import Dict

import numpy as np

from . import _lib
from ._lib import check


def read_dipole(path) -> np.ndarray:
    """mu in the piece layout [alpha][v_I][v_F] (what upload_dipole takes)."""
    L = _lib.load()
    nc = C.c_int(0)
    check(L.gain_trove_read_dipole(str(path).encode(), C.byref(nc), None))
    out = np.empty((3, nc.value, nc.value))
    check(L.gain_trove_read_dipole(str(path).encode(), C.byref(nc),
                                   out.ctypes.data))
    return out


def read_basis(path, path_j0, J: int, sym_degen) -> Dict[str, object]:
    """Basis arrays of one J after the J=0 correlation: kstart, kblock, tau,
    vib, icontr, ijterms, N and repres (lists per irrep), ntot, matsize."""
    L = _lib.load()
    nrep = len(sym_degen)
    sd = np.ascontiguousarray(sym_degen, np.int32)
    h = C.c_void_p()
    check(L.gain_trove_read_basis(str(path).encode(), str(path_j0).encode(),
                                  J, nrep, sd.ctypes.data, C.byref(h)))
    try:
        nprim, msc, mat = C.c_int(), C.c_int(), C.c_int()
        ntot = np.zeros(nrep, np.int32)
        check(L.gain_trove_basis_sizes(h, C.byref(nprim), C.byref(msc),
                                       C.byref(mat), ntot.ctypes.data))
        n, m = nprim.value, msc.value

        def get(name, size, dtype=np.int32, g=0):
            a = np.zeros(size, dtype)
            check(L.gain_trove_basis_get(h, name.encode(), g, a.ctypes.data))
            return a
        out = {"nprim": n, "maxsymcoeffs": m, "matsize": mat.value,
               "ntot": [int(x) for x in ntot],
               "kstart": get("kstart", J + 1), "kblock": get("kblock", J + 1),
               "tau": get("tau", n), "vib": get("vib", n),
               "icontr": get("icontr", 2 * n),
               "ijterms": get("ijterms", m * nrep)}
        out["N"] = [get("N", m, g=g) for g in range(nrep)]
        out["repres"] = [get("repres", int(ntot[g]) * int(sd[g]) * mat.value,
                             np.float64, g) for g in range(nrep)]
        return out
    finally:
        L.gain_trove_basis_free(h)


def read_states(path, gamma: int, params) -> Dict[str, np.ndarray]:
    """States of one descr file that pass the state filter: irec, ilevel
    (0-based), energy, and rec_len."""
    from .engine import params_struct
    L = _lib.load()
    p = params_struct(params)
    n, rl = C.c_int(0), C.c_int(0)
    check(L.gain_trove_read_states(str(path).encode(), gamma, C.byref(p),
                                   C.byref(n), C.byref(rl), None, None, None))
    irec = np.zeros(n.value, np.int32)
    ilevel = np.zeros(n.value, np.int32)
    energy = np.zeros(n.value)
    check(L.gain_trove_read_states(str(path).encode(), gamma, C.byref(p),
                                   C.byref(n), C.byref(rl), irec.ctypes.data,
                                   ilevel.ctypes.data, energy.ctypes.data))
    return {"irec": irec, "ilevel": ilevel, "energy": energy,
            "rec_len": rl.value}


def read_records(path, irec, rec_len: int) -> np.ndarray:
    L = _lib.load()
    ir = np.ascontiguousarray(irec, np.int32)
    out = np.zeros((ir.shape[0], rec_len))
    check(L.gain_trove_read_records(str(path).encode(), ir.shape[0],
                                    ir.ctypes.data, rec_len,
                                    out.ctypes.data))
    return out
