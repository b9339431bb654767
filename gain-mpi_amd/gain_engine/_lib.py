"""ctypes binding of libgain_engine.so (the C ABI in include/gain_engine.h).

The library is built in-tree by `make -C gain-mpi_amd`; there is no CPU
fallback: loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# GAIN_ENGINE_LIB selects another build of the same sources (the host
# sanitizer build, libgain_engine_asan.so); it is loaded as is, without torch
LIB_OVERRIDE = os.environ.get("GAIN_ENGINE_LIB")
LIB_PATH = LIB_OVERRIDE or os.path.join(_HERE, "libgain_engine.so")
MAX_IRREPS = 16

OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_NOMEM, ERR_CONFIG = range(6)
PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_BF16X4 = 0, 1, 2, 3
PREC_I8X5, PREC_I8X6, PREC_I8X7 = 4, 5, 6
PREC_I8CRT11, PREC_I8CRT12, PREC_I8CRT14 = 7, 8, 9

# Every symbol include/gain_engine.h declares.
EXPORTS = (
    "gain_last_error", "gain_version", "gain_create", "gain_destroy",
    "gain_get_threej", "gain_get_stream", "gain_synchronize",
    "gain_set_params", "gain_get_gamma_pair", "gain_upload_dipole",
    "gain_upload_basis", "gain_upload_inflation", "gain_upload_block",
    "gain_bind_block_device", "gain_half_ls", "gain_plain_lines",
    "gain_batched_lines", "gain_set_timing", "gain_get_timing",
    "gain_copy_symmetry_constants", "gain_copy_jmax_constant",
    "gain_copy_dipole_constant", "gain_transform_vector_primitive",
    "gain_compute_half_linestrength", "gain_compute_half_linestrength_rotsym",
    "gain_ddot", "gain_adjoint_dipole", "gain_adjoint_dipole_ls",
    "gain_rotsym_functions",
    "gain_rotsym_wigner",
    "gain_set_dipole_combination", "gain_dipole_coeff_grad",
    "gain_set_tuning", "gain_get_tuning", "gain_tuning_report",
    "gain_mfma_peak", "gain_hbm_read_peak",
    "gain_crt_moduli", "gain_crt_beta", "gain_crt_reconstruct",
    "gain_sweep_create", "gain_format_line", "gain_write_transitions",
    "gain_set_block_levels", "gain_trove_read_dipole", "gain_trove_read_basis",
    "gain_trove_basis_free", "gain_trove_basis_sizes", "gain_trove_basis_get",
    "gain_trove_upload_basis", "gain_trove_read_states",
    "gain_trove_read_records", "gain_trove_load",
    "gain_sweep_run", "gain_sweep_destroy", "gain_plain_lines_noalloc",
    "gain_upload_basis_rotsym", "gain_upload_wigner",
)


class GainError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gain_engine error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("nrepres", C.c_int), ("maxdeg", C.c_int), ("maxJ", C.c_int),
                ("reduced", C.c_int), ("symmetry_type", C.c_int),
                ("sym_degen", C.c_int * MAX_IRREPS),
                ("isym_do", C.c_int * MAX_IRREPS),
                ("isym_pairs", C.c_int * MAX_IRREPS),
                ("gns", C.c_double * MAX_IRREPS),
                ("erange_lower", C.c_double * 2),
                ("erange_upper", C.c_double * 2),
                ("freq_window", C.c_double * 2), ("zpe", C.c_double),
                ("temperature", C.c_double), ("partition", C.c_double),
                ("threshold", C.c_double)]


class Lines(C.Structure):
    _fields_ = [("ls", C.c_void_p), ("S", C.c_void_p), ("A", C.c_void_p),
                ("I", C.c_void_p), ("flags", C.c_void_p),
                ("on_device", C.c_int)]


class OutputOpts(C.Structure):
    _fields_ = [("full_line", C.c_int), ("compute_intensity", C.c_int),
                ("batch", C.c_int), ("precision", C.c_int)]


class Task(C.Structure):
    _fields_ = [("jindI", C.c_int), ("gammaI", C.c_int),
                ("levelsI", C.c_void_p), ("nI", C.c_int),
                ("jindF", C.c_int), ("gammaF", C.c_int), ("out", Lines)]


_lib = None

_vp, _ip, _dp = C.c_void_p, C.c_int, C.c_double


def load():
    """Load the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: torch ships its own libamdhip64.so.7; when
    # torch is importable it is loaded first so the engine binds to the same
    # runtime (device pointers are then shared with torch tensors).
    if not LIB_OVERRIDE:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C gain-mpi_amd` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    sig = {
        "gain_last_error": (C.c_char_p, []),
        "gain_version": (C.c_char_p, []),
        "gain_create": (_ip, [C.POINTER(_vp), _ip, _ip, _ip, _ip]),
        "gain_destroy": (_ip, [_vp]),
        "gain_get_threej": (_ip, [_vp, _vp]),
        "gain_get_stream": (_ip, [_vp, C.POINTER(_vp)]),
        "gain_synchronize": (_ip, [_vp]),
        "gain_set_params": (_ip, [_vp, C.POINTER(Params)]),
        "gain_get_gamma_pair": (_ip, [_vp, _ip, C.POINTER(_ip)]),
        "gain_upload_dipole": (_ip, [_vp, _vp, _ip, _ip, _ip]),
        "gain_upload_basis": (_ip, [_vp, _ip, _ip, _ip, _vp, _vp, _vp, _vp]),
        "gain_upload_inflation": (_ip, [_vp, _ip, _ip, _vp, _vp, _ip, _vp,
                                        _vp, _vp, _vp]),
        "gain_upload_basis_rotsym": (_ip, [_vp, _ip, _ip, _ip, _ip, _ip, _vp,
                                           _vp, _vp]),
        "gain_upload_wigner": (_ip, [_vp, _ip, _ip, _ip, _ip, _ip, _vp]),
        "gain_upload_block": (_ip, [_vp, _ip, _ip, _ip, _ip, _vp, _vp]),
        "gain_bind_block_device": (_ip, [_vp, _ip, _ip, _ip, _ip,
                                         C.c_int64, _vp, _vp]),
        "gain_half_ls": (_ip, [_vp, _ip, _ip, _ip, _vp, _ip, _vp]),
        "gain_plain_lines": (_ip, [_vp, _ip, _ip, _ip, _vp, _ip, _ip,
                                   C.POINTER(Lines)]),
        "gain_batched_lines": (_ip, [_vp, _ip, _ip, _vp, _ip, _ip, _ip, _ip,
                                     C.POINTER(Lines)]),
        "gain_plain_lines_noalloc": (_ip, [_vp, _ip, _ip, _ip, _vp, _ip, _ip,
                                           C.POINTER(Lines)]),
        "gain_set_timing": (_ip, [_vp, _ip]),
        "gain_set_block_levels": (_ip, [_vp, _ip, _ip, _vp]),
        "gain_trove_read_dipole": (_ip, [C.c_char_p, C.POINTER(_ip), _vp]),
        "gain_trove_read_basis": (_ip, [C.c_char_p, C.c_char_p, _ip, _ip,
                                        _vp, C.POINTER(_vp)]),
        "gain_trove_basis_free": (_ip, [_vp]),
        "gain_trove_basis_sizes": (_ip, [_vp, C.POINTER(_ip), C.POINTER(_ip),
                                         C.POINTER(_ip), _vp]),
        "gain_trove_basis_get": (_ip, [_vp, C.c_char_p, _ip, _vp]),
        "gain_trove_upload_basis": (_ip, [_vp, _ip, _vp]),
        "gain_trove_read_states": (_ip, [C.c_char_p, _ip, C.POINTER(Params),
                                         C.POINTER(_ip), C.POINTER(_ip), _vp,
                                         _vp, _vp]),
        "gain_trove_read_records": (_ip, [C.c_char_p, _ip, _vp, _ip, _vp]),
        "gain_trove_load": (_ip, [_vp, C.c_char_p, _vp, _ip]),
        "gain_format_line": (_ip, [C.c_char_p, _ip, _dp, _ip, _ip, _ip, _ip,
                                   _ip, _ip, _dp, _ip, _dp, _vp, _ip]),
        "gain_write_transitions": (_ip, [_vp, C.POINTER(OutputOpts),
                                         C.c_char_p,
                                         C.POINTER(C.c_int64)]),
        "gain_get_timing": (_ip, [_vp, C.c_char_p, C.POINTER(_dp),
                                  C.POINTER(C.c_int64), _ip]),
        "gain_set_tuning": (_ip, [_vp, C.c_char_p, _ip]),
        "gain_get_tuning": (_ip, [_vp, C.c_char_p,
                                 C.POINTER(C.c_longlong)]),
        "gain_tuning_report": (_ip, [_vp, C.c_char_p, _ip]),
        "gain_mfma_peak": (_ip, [_vp, C.c_char_p, C.POINTER(_dp)]),
        "gain_hbm_read_peak": (_ip, [_vp, C.POINTER(_dp)]),
        "gain_crt_moduli": (_ip, [C.c_int, C.POINTER(C.c_int)]),
        "gain_crt_beta": (_ip, [C.c_int, C.c_longlong, C.POINTER(C.c_int),
                                C.POINTER(C.c_int)]),
        "gain_crt_reconstruct": (_ip, [C.c_int, C.POINTER(C.c_uint),
                                       C.c_longlong, C.POINTER(_dp)]),
        "gain_sweep_create": (_ip, [_vp, _vp, _ip, _ip, C.POINTER(_vp)]),
        "gain_sweep_run": (_ip, [_vp]),
        "gain_sweep_destroy": (_ip, [_vp]),
        "gain_copy_symmetry_constants": (_ip, [_ip, _ip]),
        "gain_copy_jmax_constant": (_ip, [_ip]),
        "gain_copy_dipole_constant": (_ip, [_ip]),
        "gain_transform_vector_primitive": (_ip, [_ip] * 6 + [_vp] * 7),
        "gain_compute_half_linestrength": (
            _ip, [_ip] * 5 + [_vp] * 4 + [_ip] * 6 + [_vp] * 5 + [_vp]),
        "gain_compute_half_linestrength_rotsym": (
            _ip, [_ip] * 5 + [_vp] * 6 + [_ip] * 3 + [_vp] * 5),
        "gain_ddot": (_ip, [_ip, _vp, _vp, _vp, _vp]),
        "gain_rotsym_functions": (_ip, [C.c_char_p, _ip] + [_vp] * 6),
        "gain_rotsym_wigner": (_ip, [C.c_char_p, _ip, _ip] + [_vp] * 4),
        "gain_adjoint_dipole": (_ip, [_vp, _ip, _ip, _vp, _ip, _ip, _ip, _vp,
                                      _vp, _ip, _ip]),
        "gain_adjoint_dipole_ls": (_ip, [_vp, _ip, _ip, _vp, _ip, _ip, _ip,
                                         _vp, _vp, _vp, _vp, _ip, _ip]),
        "gain_set_dipole_combination": (_ip, [_vp, _ip, _vp, _vp, _ip]),
        "gain_dipole_coeff_grad": (_ip, [_vp, _ip, _vp, _vp, _vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int):
    if rc != OK:
        msg = load().gain_last_error().decode(errors="replace")
        raise GainError(rc, msg)
