"""Host-side front end of the line-strength engine (Python over the C ABI).

`Engine` plays the role of the reference's GpuManager
(BaseClasses/GpuManager.h:136-172): it owns one device context, receives
the dipole, the per-J basis and inflation data and the state blocks, and
runs the hot path per initial state (plain FP64 GEMV drop-in) or per batch
of initial states (GEMM engine).
"""
from __future__ import annotations

import ctypes as C
import sys
from typing import Dict, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check

_i32 = np.int32
_f64 = np.float64


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def params_struct(p) -> _lib.Params:
    s = _lib.Params()
    s.nrepres = p.nrepres
    s.maxdeg = p.maxdeg
    s.maxJ = p.maxJ
    s.reduced = int(bool(p.reduced))
    s.symmetry_type = p.symmetry_type
    for g in range(p.nrepres):
        s.sym_degen[g] = p.sym_degen[g]
        s.isym_do[g] = int(p.isym_do[g])
        s.isym_pairs[g] = p.isym_pairs[g]
        s.gns[g] = p.gns[g]
    s.erange_lower[:] = list(p.erange_lower)
    s.erange_upper[:] = list(p.erange_upper)
    s.freq_window[:] = list(p.freq_window)
    s.zpe, s.temperature = p.zpe, p.temperature
    s.partition, s.threshold = p.partition, p.threshold
    return s


class Engine:
    def __init__(self, device: int, jmax: int, nrepres: int, maxdeg: int):
        L = _lib.load()
        h = C.c_void_p()
        check(L.gain_create(C.byref(h), device, jmax, nrepres, maxdeg))
        self._L, self._h = L, h
        self.device = device
        self.jmax, self.nrepres, self.maxdeg = jmax, nrepres, maxdeg
        self._keep = []
        self._xs = None
        self.blocks: Dict = {}

    def wait_for_torch(self):
        """Order the engine's stream after the work torch has queued on its
        current stream, so device tensors torch is still writing are
        complete before engine kernels read or overwrite them (no host
        synchronisation: the engine's stream waits on a torch event)."""
        torch = sys.modules.get("torch")
        if torch is None or not torch.cuda.is_initialized():
            return
        if self._xs is None:
            self._xs = torch.cuda.ExternalStream(self.stream(),
                                                 device=self.device)
        self._xs.wait_stream(torch.cuda.current_stream(self.device))

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if self._h:
            check(self._L.gain_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ setup
    def threej(self) -> np.ndarray:
        t = np.zeros((self.jmax + 1) ** 2 * 9)
        check(self._L.gain_get_threej(self._h, _ptr(t)))
        return t

    def stream(self) -> int:
        s = C.c_void_p()
        check(self._L.gain_get_stream(self._h, C.byref(s)))
        return s.value or 0

    def synchronize(self):
        check(self._L.gain_synchronize(self._h))

    def set_params(self, p):
        s = params_struct(p)
        check(self._L.gain_set_params(self._h, C.byref(s)))

    def gamma_pair(self, g: int) -> int:
        out = C.c_int()
        check(self._L.gain_get_gamma_pair(self._h, g, C.byref(out)))
        return out.value

    def upload_dipole(self, mu: np.ndarray, start: int = 0, end=None):
        """mu[3][nc(v_I)][w(v_F)]: the whole dipole (w = nc), or the piece
        of final columns [start, end) in the reference piece layout."""
        mu = np.ascontiguousarray(mu, _f64)
        nc = mu.shape[1]
        end = start + mu.shape[2] if end is None else end
        if mu.shape != (3, nc, end - start):
            raise ValueError(f"dipole piece of shape {mu.shape} is not "
                             f"[3][{nc}][{end - start}]")
        check(self._L.gain_upload_dipole(self._h, _ptr(mu), nc, start, end))

    def upload_basis(self, jind: int, b):
        if getattr(b, "rotsym", False):
            arrs = [np.ascontiguousarray(getattr(b, n), _i32)
                    for n in ("level", "deg", "vib")]
            check(self._L.gain_upload_basis_rotsym(
                self._h, jind, b.J, b.nprim, b.nlevels, b.rmd,
                *[_ptr(a) for a in arrs]))
            return
        arrs = [np.ascontiguousarray(getattr(b, n), _i32)
                for n in ("kstart", "kblock", "tau", "vib")]
        check(self._L.gain_upload_basis(self._h, jind, b.J, b.nprim,
                                        *[_ptr(a) for a in arrs]))

    def upload_wigner(self, jindI: int, jindF: int, nlevelsI: int,
                      nlevelsF: int, maxdeg: int, rot: np.ndarray):
        r = np.ascontiguousarray(rot, _f64)
        check(self._L.gain_upload_wigner(self._h, jindI, jindF, nlevelsI,
                                         nlevelsF, maxdeg, _ptr(r)))

    def upload_inflation(self, jind: int, b):
        nrep = self.nrepres
        ij = np.ascontiguousarray(b.ijterms, _i32)
        ic = np.ascontiguousarray(b.icontr, _i32)
        ntot = np.array(list(b.ntot) + [0] * (nrep - len(b.ntot)), _i32)
        sdeg = np.array(list(b.sdeg) + [1] * (nrep - len(b.sdeg)), _i32)
        Ns = [np.ascontiguousarray(b.N[g], _i32) for g in range(nrep)]
        Rs = [np.ascontiguousarray(b.repres[g], _f64) if b.repres[g].size
              else np.zeros(1) for g in range(nrep)]
        Np = (C.c_void_p * nrep)(*[_ptr(a) for a in Ns])
        Rp = (C.c_void_p * nrep)(*[_ptr(a) for a in Rs])
        check(self._L.gain_upload_inflation(
            self._h, jind, b.maxsymcoeffs, _ptr(ij), _ptr(ic), b.matsize,
            _ptr(ntot), _ptr(sdeg), C.cast(Np, C.c_void_p),
            C.cast(Rp, C.c_void_p)))

    def upload_block(self, jind, gamma, energies, vectors=None, rec_len=None):
        e = np.ascontiguousarray(energies, _f64)
        if vectors is not None:
            v = np.ascontiguousarray(vectors, _f64)
            rl = v.shape[1]
            vp = _ptr(v)
        else:
            rl, vp = int(rec_len), None
        check(self._L.gain_upload_block(self._h, jind, gamma, e.shape[0], rl,
                                        _ptr(e), vp))
        self.blocks[(jind, gamma)] = (e.shape[0], rl)

    def bind_block_device(self, jind, gamma, energies, dev_ptr: int,
                          rec_len: int, ld: int):
        e = np.ascontiguousarray(energies, _f64)
        self.wait_for_torch()
        check(self._L.gain_bind_block_device(self._h, jind, gamma,
                                             e.shape[0], rec_len, ld,
                                             _ptr(e), dev_ptr))
        self.blocks[(jind, gamma)] = (e.shape[0], rec_len)

    @classmethod
    def from_problem(cls, prob, device: int = 0, upload_vectors=True,
                     device_blocks: Optional[Dict] = None,
                     dipole_pieces=None) -> "Engine":
        """Upload a synth.Problem.  `device_blocks` maps (jind, gamma) to a
        (device pointer, ld) of records already in HBM (bound, not copied);
        `dipole_pieces` lists (start, end) ranges of final columns to upload
        the dipole as, in that order, instead of whole."""
        eng = cls(device, prob.jmax, prob.params.nrepres, prob.params.maxdeg)
        eng.set_params(prob.params)
        if dipole_pieces is None:
            eng.upload_dipole(prob.mu)
        else:
            for a, b in dipole_pieces:
                eng.upload_dipole(prob.mu[:, :, a:b], a, b)
        for jind, b in enumerate(prob.bases):
            eng.upload_basis(jind, b)
            eng.upload_inflation(jind, b)
        for jI, bI in enumerate(prob.bases):
            for jF, bF in enumerate(prob.bases):
                if getattr(bI, "rotsym", False) and bI.wig and \
                        (bF.J - bI.J + 1) in bI.wig:
                    eng.upload_wigner(jI, jF, bI.nlevels, bF.nlevels, bI.rmd,
                                      bI.wig[bF.J - bI.J + 1])
        for key, blk in prob.blocks.items():
            if device_blocks and key in device_blocks:
                ptr, ld = device_blocks[key]
                eng.bind_block_device(key[0], key[1], blk.energy, ptr,
                                      blk.rec_len, ld)
            elif upload_vectors and blk.vectors is not None:
                eng.upload_block(key[0], key[1], blk.energy, blk.vectors)
            else:
                eng.upload_block(key[0], key[1], blk.energy,
                                 rec_len=blk.rec_len)
        return eng

    # ------------------------------------------------------------ hot path
    def half_ls(self, jindI, gammaI, idegI, cI, jindF, nprimF) -> np.ndarray:
        c = np.ascontiguousarray(cI, _f64)
        h = np.zeros(nprimF)
        check(self._L.gain_half_ls(self._h, jindI, gammaI, idegI, _ptr(c),
                                   jindF, _ptr(h)))
        return h

    def _host_lines(self, nI, nF, want_ls=True):
        md2 = self.maxdeg * self.maxdeg
        out = {"S": np.zeros((nI, nF)), "A": np.zeros((nI, nF)),
               "I": np.zeros((nI, nF)),
               "flags": np.zeros((nI, nF), np.uint8)}
        if want_ls:
            out["ls"] = np.zeros((nI, nF, md2))
        s = _lib.Lines()
        for k in ("ls", "S", "A", "I", "flags"):
            setattr(s, k, _ptr(out[k]) if k in out else None)
        s.on_device = 0
        return out, s

    def plain_lines(self, jindI, gammaI, levelI, jindF, gammaF, cI=None,
                    want_ls=True):
        nF = self.blocks[(jindF, gammaF)][0]
        out, s = self._host_lines(1, nF, want_ls)
        c = None if cI is None else np.ascontiguousarray(cI, _f64)
        check(self._L.gain_plain_lines(self._h, jindI, gammaI, levelI,
                                       None if c is None else _ptr(c), jindF,
                                       gammaF, C.byref(s)))
        return {k: v[0] for k, v in out.items()}

    def batched_lines(self, jindI, gammaI, levels: Sequence[int], jindF,
                      gammaF, precision=_lib.PREC_FP64, want_ls=True):
        lv = np.ascontiguousarray(levels, _i32)
        nF = self.blocks[(jindF, gammaF)][0]
        out, s = self._host_lines(lv.shape[0], nF, want_ls)
        check(self._L.gain_batched_lines(self._h, jindI, gammaI, _ptr(lv),
                                         lv.shape[0], jindF, gammaF,
                                         precision, C.byref(s)))
        return out

    def lines_device(self, jindI, gammaI, levels, jindF, gammaF, ptrs: Dict,
                     precision=_lib.PREC_FP64, plain=False, cI=None):
        """Write results into caller device buffers (dict of name -> ptr)."""
        self.wait_for_torch()
        s = _lib.Lines()
        for k in ("ls", "S", "A", "I", "flags"):
            setattr(s, k, ptrs.get(k))
        s.on_device = 1
        if plain:
            c = None if cI is None else np.ascontiguousarray(cI, _f64)
            check(self._L.gain_plain_lines(
                self._h, jindI, gammaI, int(levels[0]),
                None if c is None else _ptr(c), jindF, gammaF, C.byref(s)))
        else:
            lv = np.ascontiguousarray(levels, _i32)
            check(self._L.gain_batched_lines(self._h, jindI, gammaI,
                                             _ptr(lv), lv.shape[0], jindF,
                                             gammaF, precision, C.byref(s)))

    # ------------------------------------------------------------ timing
    def set_timing(self, on: bool):
        check(self._L.gain_set_timing(self._h, int(on)))

    def timing(self, name: str, reset=False):
        ms, n = C.c_double(), C.c_int64()
        check(self._L.gain_get_timing(self._h, name.encode(), C.byref(ms),
                                      C.byref(n), int(reset)))
        return ms.value, n.value

    def set_tuning(self, key: str, value: int):
        check(self._L.gain_set_tuning(self._h, key.encode(), int(value)))

    def write_transitions(self, path, full_line=False, intensity=True,
                          batch=64, precision=_lib.PREC_FP64) -> int:
        """Every line of the uploaded blocks in the reference's output format
        (main.cpp:238-338, Output.cpp:52-112); returns the line count."""
        o = _lib.OutputOpts(int(full_line), int(intensity), int(batch),
                            int(precision))
        n = C.c_int64(0)
        check(self._L.gain_write_transitions(
            self._h, C.byref(o), None if path is None else str(path).encode(),
            C.byref(n)))
        return int(n.value)

    def set_block_levels(self, jind, gamma, levels):
        lv = None if levels is None else np.ascontiguousarray(levels, _i32)
        check(self._L.gain_set_block_levels(
            self._h, jind, gamma, None if lv is None else _ptr(lv)))

    def trove_load(self, directory, jvals):
        """Dipole, bases, states and records from TROVE checkpoint files in
        `directory` for the J list `jvals` (J index = position)."""
        jv = np.ascontiguousarray(jvals, _i32)
        check(self._L.gain_trove_load(self._h, str(directory).encode(),
                                      _ptr(jv), jv.shape[0]))

    def get_tuning(self, key: str) -> int:
        v = C.c_longlong(0)
        check(self._L.gain_get_tuning(self._h, key.encode(), C.byref(v)))
        return int(v.value)

    def mfma_peak(self, kind: str) -> float:
        """Measured register-only MFMA throughput ("i8" or "f64"), ops/s."""
        v = C.c_double(0.0)
        check(self._L.gain_mfma_peak(self._h, kind.encode(), C.byref(v)))
        return float(v.value)

    def hbm_read_peak(self) -> float:
        """Measured read-only HBM streaming rate, bytes/s."""
        v = C.c_double(0.0)
        check(self._L.gain_hbm_read_peak(self._h, C.byref(v)))
        return float(v.value)

    def tuning_report(self) -> str:
        buf = C.create_string_buffer(1 << 16)
        check(self._L.gain_tuning_report(self._h, buf, 1 << 16))
        return buf.value.decode()


def _attach_adjoint():
    """Adjoint / dipole-parameterisation methods of Engine."""

    def adjoint(self, jindI, gammaI, levels, jindF, gammaF, dS, dmu,
                on_device=True, accumulate=False, ls=None, flags=None):
        """dL/dmu for one batch given dS = dL/dS ([nI][nF]); dS and dmu are
        device pointers (ints) when on_device, else numpy arrays.  With ls
        and flags (the FP64 batched forward's gain_lines.ls and .flags for
        this batch, same kind of memory) the forward is not recomputed."""
        lv = np.ascontiguousarray(levels, _i32)
        if on_device:
            self.wait_for_torch()
            p_dS, p_dmu = int(dS), int(dmu)
            p_ls = None if ls is None else int(ls)
            p_fl = None if flags is None else int(flags)
        else:
            dS = np.ascontiguousarray(dS, _f64)
            p_dS, p_dmu = _ptr(dS), _ptr(dmu)
            if ls is not None:
                ls = np.ascontiguousarray(ls, _f64)
                flags = np.ascontiguousarray(flags, np.uint8)
            p_ls = None if ls is None else _ptr(ls)
            p_fl = None if flags is None else _ptr(flags)
        if p_ls is None and p_fl is None:
            check(self._L.gain_adjoint_dipole(self._h, jindI, gammaI,
                                              _ptr(lv), lv.shape[0], jindF,
                                              gammaF, p_dS, p_dmu,
                                              int(on_device),
                                              int(accumulate)))
            return
        check(self._L.gain_adjoint_dipole_ls(self._h, jindI, gammaI,
                                             _ptr(lv), lv.shape[0], jindF,
                                             gammaF, p_dS, p_ls, p_fl, p_dmu,
                                             int(on_device),
                                             int(accumulate)))

    def set_dipole_combination(self, coef, M_ptr: int, nc: int):
        cf = np.ascontiguousarray(coef, _f64)
        self.wait_for_torch()
        check(self._L.gain_set_dipole_combination(self._h, cf.shape[0],
                                                  _ptr(cf), M_ptr, nc))

    def dipole_coeff_grad(self, P: int, M_ptr: int, dmu_ptr: int):
        dc = np.zeros(P)
        self.wait_for_torch()
        check(self._L.gain_dipole_coeff_grad(self._h, P, M_ptr, dmu_ptr,
                                             _ptr(dc)))
        return dc

    Engine.adjoint = adjoint
    Engine.set_dipole_combination = set_dipole_combination
    Engine.dipole_coeff_grad = dipole_coeff_grad


_attach_adjoint()


class Sweep:
    """A captured mixed-J sweep (gain_sweep_*): tasks are tuples
    (jindI, gammaI, levels, jindF, gammaF, {name: device_ptr}) whose
    outputs are device buffers written by every run()."""

    def __init__(self, eng: Engine, tasks, precision=_lib.PREC_FP64):
        self._eng = eng
        self._keep = []
        self._h = None
        arr = (_lib.Task * len(tasks))()
        for k, (jI, gI, levels, jF, gF, ptrs) in enumerate(tasks):
            lv = np.ascontiguousarray(levels, _i32)
            self._keep.append(lv)
            t = arr[k]
            t.jindI, t.gammaI, t.jindF, t.gammaF = jI, gI, jF, gF
            t.levelsI, t.nI = _ptr(lv), lv.shape[0]
            for name in ("ls", "S", "A", "I", "flags"):
                setattr(t.out, name, ptrs.get(name))
            t.out.on_device = 1
        h = C.c_void_p()
        check(eng._L.gain_sweep_create(eng._h, C.cast(arr, C.c_void_p),
                                       len(tasks), precision, C.byref(h)))
        self._h = h

    def run(self):
        self._eng.wait_for_torch()
        check(self._eng._L.gain_sweep_run(self._h))

    def close(self):
        if self._h:
            check(self._eng._L.gain_sweep_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
