"""The host-side tests of the library's C++ host code (TROVE checkpoint
readers including truncated and bogus files, the ROTSYM table generator,
the line formats, the ABI error paths) run again against the sanitizer
build (AddressSanitizer + UndefinedBehaviorSanitizer on the host code,
`make -C gain-mpi_amd asan`) in a child process with the ASan runtime
preloaded; any sanitizer report fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN_LIB = os.path.join(ROOT, "gain-mpi_amd", "gain_engine",
                        "libgain_engine_asan.so")
HOST_TESTS = ["tests/test_trove_io.py", "tests/test_rotsym.py",
              "tests/test_writer_format.py", "tests/test_abi.py"]


def _asan_runtime():
    try:
        out = subprocess.run(["/opt/rocm/llvm/bin/clang++",
                              "-print-file-name=libclang_rt.asan-x86_64.so"],
                             capture_output=True, text=True, timeout=60)
    except (OSError, subprocess.TimeoutExpired):
        return None
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_host_code_under_address_and_ub_sanitizers(tmp_path):
    rt = _asan_runtime()
    if not os.path.exists(ASAN_LIB) or rt is None:
        pytest.skip("sanitizer build absent (make -C gain-mpi_amd asan)")
    log = tmp_path / "san"
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": rt,
        "ASAN_OPTIONS": f"detect_leaks=0:halt_on_error=1:log_path={log}a",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1:"
                         f"log_path={log}u",
        "GAIN_ENGINE_LIB": ASAN_LIB,
    })
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x",
                        "-p", "no:cacheprovider", *HOST_TESTS],
                       cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=1200)
    reports = [p for p in tmp_path.iterdir() if p.name.startswith("san")]
    text = "".join(p.read_text() for p in reports)
    assert not reports, text[:4000]
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert " passed" in r.stdout
