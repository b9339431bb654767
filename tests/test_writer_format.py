"""Line format of the transition writer (Output.cpp:99-111), host-only:
gain_format_line against the reference's printf formats applied to the
same numbers, and against the literal pinned for the oracle's formatter."""
import ctypes as C

import numpy as np
import pytest

from gain_engine import _lib


def _fmt(nu, lF, jF, gF, lI, jI, gI, A, inten=None, ls=None):
    L = _lib.load()
    buf = C.create_string_buffer(4096)
    arr = None if ls is None else np.ascontiguousarray(ls, np.float64)
    rc = L.gain_format_line(buf, 4096, nu, lF, jF, gF, lI, jI, gI, A,
                            0 if inten is None else 1,
                            0.0 if inten is None else inten,
                            None if arr is None else arr.ctypes.data,
                            -1 if arr is None else len(arr))
    _lib.check(rc)
    return buf.value.decode()


def _reference_format(nu, lF, jF, gF, lI, jI, gI, A, inten=None, ls=None):
    s = "%12.6f %8d %4d %4d <- %8d %4d %4d %16.9E " % (
        nu, lF + 1, jF, gF + 1, lI + 1, jI, gI + 1, A)
    if inten is not None:
        s += "%16.9E " % inten
    if ls is not None:
        s += "[ %s ] " % "".join(" %16.9E" % v for v in ls)
    return s + " || \n"


def test_format_line_literal():
    A = 3.13618923e-7 * 1 * 0.25 * 1234.5 ** 3
    txt = _fmt(1234.5, 9, 1, 1, 4, 0, 0, A, ls=[0.5])
    assert txt.startswith(" 1234.500000       10    1    2 <-        5    0"
                          "    1 ")
    assert f"{A:16.9E} " in txt
    assert txt.endswith("[   5.000000000E-01 ]  || \n")


@pytest.mark.parametrize("with_I,nls", [(False, None), (True, None),
                                        (True, 0), (True, 1), (False, 4),
                                        (True, 9)])
def test_format_line_matches_reference_printf(with_I, nls):
    rng = np.random.default_rng(nls or 0)
    for _ in range(200):
        nu = float(rng.uniform(1e-3, 2e4))
        A = float(10 ** rng.uniform(-30, 3)) * rng.choice([1, -1])
        inten = float(10 ** rng.uniform(-40, -15)) if with_I else None
        ls = None if nls is None else rng.standard_normal(nls) * 1e-3
        args = (nu, int(rng.integers(0, 10 ** 6)), int(rng.integers(0, 80)),
                int(rng.integers(0, 16)), int(rng.integers(0, 10 ** 6)),
                int(rng.integers(0, 80)), int(rng.integers(0, 16)), A, inten,
                ls)
        assert _fmt(*args) == _reference_format(*args)


def test_format_line_rejects_small_buffer():
    L = _lib.load()
    buf = C.create_string_buffer(16)
    rc = L.gain_format_line(buf, 16, 1.0, 0, 0, 0, 0, 0, 0, 1.0, 0, 0.0,
                            None, -1)
    assert rc == 1
    assert b"too small" in L.gain_last_error()
