"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/gain_engine.h declares, and argument errors come back as
status codes (no HIP call is made on these paths)."""
import ctypes as C
import os
import re

from gain_engine import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "gain_engine.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gain_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    names = declared()
    assert len(names) >= 20
    assert sorted(_lib.EXPORTS) == names


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    for n in declared():
        assert hasattr(L, n), n


def test_version_and_error_codes_without_gpu():
    L = _lib.load()
    assert b"gfx950" in L.gain_version()
    assert L.gain_copy_jmax_constant(-1) == _lib.ERR_ARG
    assert b"jmax" in L.gain_last_error()
    assert L.gain_copy_dipole_constant(0) == _lib.ERR_ARG
    assert L.gain_copy_symmetry_constants(0, 1) == _lib.ERR_ARG
    assert L.gain_copy_jmax_constant(5) == _lib.OK
    assert L.gain_last_error() == b""
    # null context / null output are rejected before any device call
    assert L.gain_destroy(None) == _lib.OK
    assert L.gain_set_timing(None, 1) == _lib.ERR_ARG
    assert L.gain_create(None, 0, 1, 1, 1) == _lib.ERR_ARG


def test_python_front_end_imports():
    import importlib
    for m in ("gain_engine.synth", "gain_engine.engine", "gain_engine.workload",
              "gain_engine.autograd"):
        importlib.import_module(m)
