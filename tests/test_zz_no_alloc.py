"""The allocation-free plain drop-in (gain_plain_lines_noalloc) through the
C ABI: two calls in a row, the second on a smaller input, each bit-identical
to gain_plain_lines on the same inputs (same launches and arguments), no
change of the device's free memory across the calls, and the argument
errors (host outputs, buffers above the array's cap)."""
import ctypes as C

import numpy as np
import pytest

from gain_engine import _lib, synth
from gain_engine.engine import Engine

pytestmark = pytest.mark.gpu

KEYS = ("ls", "S", "A", "I", "flags")


def _problem(scheme):
    if scheme == "c3v3-full":
        # every (dI, dF) component pair: four projected vectors per state
        p = synth.Params(nrepres=3, sym_degen=[1, 1, 2], gns=[1.0] * 3,
                         isym_pairs=[1, 1, 2], maxJ=3, symmetry_type=2,
                         reduced=False)
        return synth.make_problem([1, 2, 3], 48, "c3v3", params=p)
    return synth.make_problem([1, 2, 3], 64, scheme)


def _outputs(eng, nF):
    import torch
    md2 = eng.maxdeg * eng.maxdeg
    d = {k: torch.full((nF,), -7.0, dtype=torch.float64, device="cuda:0")
         for k in ("S", "A", "I")}
    d["ls"] = torch.full((nF, md2), -7.0, dtype=torch.float64,
                         device="cuda:0")
    d["flags"] = torch.full((nF,), 255, dtype=torch.uint8, device="cuda:0")
    return d


def _lines(d):
    s = _lib.Lines()
    for k in KEYS:
        setattr(s, k, d[k].data_ptr())
    s.on_device = 1
    return s


def _noalloc(eng, jI, gI, lv, jF, gF, d, cI_ptr=None):
    L = _lib.load()
    return L.gain_plain_lines_noalloc(eng._h, jI, gI, lv, cI_ptr, jF, gF,
                                      C.byref(_lines(d)))


def _plain(eng, jI, gI, lv, jF, gF, d):
    L = _lib.load()
    _lib.check(L.gain_plain_lines(eng._h, jI, gI, lv, None, jF, gF,
                                  C.byref(_lines(d))))


def _same(a, b):
    import torch
    for k in KEYS:
        assert torch.equal(a[k], b[k]), k


@pytest.mark.parametrize("scheme", ["c2v4", "c3v3-full"])
def test_noalloc_matches_plain_two_calls(scheme):
    import torch
    prob = _problem(scheme)
    with Engine.from_problem(prob) as eng:
        g0 = 0 if scheme == "c2v4" else 2
        gF = eng.gamma_pair(g0)
        # large call first (J=3 -> J=2), then a smaller one (J=1 -> J=1)
        calls = [(2, g0, 3, 1, gF), (0, g0, 1, 0, gF)]
        # warm the runtime once (first launches of each kernel)
        jI, gI, lv, jF, gFF = calls[0]
        w = _outputs(eng, prob.blocks[(jF, gFF)].n)
        _lib.check(_noalloc(eng, jI, gI, lv + 1, jF, gFF, w))
        eng.synchronize()
        torch.cuda.synchronize()
        outs = [(_outputs(eng, prob.blocks[(c[3], c[4])].n),
                 _outputs(eng, prob.blocks[(c[3], c[4])].n)) for c in calls]
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        for (jI, gI, lv, jF, gFF), (na, _) in zip(calls, outs):
            _lib.check(_noalloc(eng, jI, gI, lv, jF, gFF, na))
        eng.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 == free0, (free0, free1)
        for (jI, gI, lv, jF, gFF), (na, pl) in zip(calls, outs):
            _plain(eng, jI, gI, lv, jF, gFF, pl)
        eng.synchronize()
        for (na, pl) in outs:
            _same(na, pl)
            assert int((na["flags"] & 1).sum()) > 0


def test_noalloc_device_record_and_errors():
    import torch
    prob = _problem("c2v4")
    with Engine.from_problem(prob) as eng:
        gF = eng.gamma_pair(0)
        nF = prob.blocks[(1, gF)].n
        lv = 5
        rec = torch.as_tensor(prob.blocks[(2, 0)].vector(lv),
                              device="cuda:0").contiguous()
        a, b = _outputs(eng, nF), _outputs(eng, nF)
        _lib.check(_noalloc(eng, 2, 0, lv, 1, gF, a, rec.data_ptr()))
        _plain(eng, 2, 0, lv, 1, gF, b)
        eng.synchronize()
        _same(a, b)
        L = _lib.load()
        # host outputs are refused
        h = {k: np.zeros(nF) for k in ("S", "A", "I")}
        s = _lib.Lines()
        s.S = h["S"].ctypes.data
        s.on_device = 0
        assert L.gain_plain_lines_noalloc(eng._h, 2, 0, lv, None, 1, gF,
                                          C.byref(s)) == _lib.ERR_ARG
        assert b"device outputs" in L.gain_last_error()
        # buffers beyond the usable part of the array are refused
        eng.set_tuning("noalloc_cap", 4096)
        assert _noalloc(eng, 2, 0, lv, 1, gF, a) == _lib.ERR_ARG
        assert b"bytes of scratch" in L.gain_last_error()
        eng.set_tuning("noalloc_cap", 0)
        c = _outputs(eng, nF)
        _lib.check(_noalloc(eng, 2, 0, lv, 1, gF, c))
        eng.synchronize()
        _same(c, b)
