"""The C ABI's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

libpxl is rebuilt with both sanitizers on its host code only (every -fsanitize= follows
-Xarch_host; the device code is built as usual), and tests/test_abi.py -- argument validation,
error codes, empty inputs, no device -- runs against that build in a subprocess with the GCC
sanitizer runtimes preloaded (ROCm's clang ships none; the instrumentation ABI is the same).  A
second subprocess hands the host-side accounting entry point a counts array one element short:
AddressSanitizer must stop it, which shows the instrumented build is the one loaded and live.
Needs no GPU.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "plenoxels_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return os.path.realpath(path) if path and os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def sanitized_lib(tmp_path_factory):
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not (os.path.exists(HIPCC) and shutil.which("gcc") and asan and ubsan):
        pytest.skip("hipcc or the GCC sanitizer runtimes are not available")
    out = str(tmp_path_factory.mktemp("asan") / "libpxl_asan.so")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           "-munsafe-fp-atomics", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize=function", "-Xarch_host", "-fno-sanitize-recover=undefined",
           "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-o", out,
           os.path.join(CSRC, "pxl_render.hip"), os.path.join(CSRC, "pxl_ops.hip"),
           os.path.join(CSRC, "pxl_accounting.cpp")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-2000:]
    env = dict(os.environ, PXL_LIBRARY=out, LD_PRELOAD=f"{asan}:{ubsan}",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    return out, env


def test_abi_tests_pass_under_asan_and_ubsan(sanitized_lib):
    out, env = sanitized_lib
    res = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_abi.py"), "-q",
                          "-p", "no:cacheprovider", "-m", "not gpu"],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    report = res.stdout[-3000:] + res.stderr[-3000:]
    assert res.returncode == 0, report
    assert "ERROR: AddressSanitizer" not in report and "runtime error:" not in report, report


PROBE = r"""
import ctypes, sys
import numpy as np
sys.path.insert(0, ROOT)
from plenoxels_amd import _lib
assert _lib.LIB_PATH.endswith("libpxl_asan.so"), _lib.LIB_PATH
L = _lib.lib()
g = _lib.Grid(); g.dims[0] = g.dims[1] = g.dims[2] = 8; g.channels = 4
m = _lib.March(); m.n_samples = 8; m.interp = _lib.INTERP_NEAREST; m.delta = 0.5; m.pd = 1.0; m.inv_pd = 1.0
o = np.zeros((16, 3), np.float32); d = np.zeros((16, 3), np.float32); d[:, 2] = 1
counts = np.zeros(15, np.int64)  # one short of the 16 rays
L.pxl_count_cells_host(ctypes.byref(g), ctypes.byref(m), o.ctypes.data, d.ctypes.data, 16, None,
                       counts.ctypes.data, 1)
print("NOT STOPPED")
"""


def test_asan_is_live_in_the_sanitized_build(sanitized_lib):
    out, env = sanitized_lib
    res = subprocess.run([sys.executable, "-c", f"ROOT = {ROOT!r}\n" + PROBE], cwd=ROOT, env=env,
                         capture_output=True, text=True, timeout=300)
    assert res.returncode != 0 and "NOT STOPPED" not in res.stdout
    assert "heap-buffer-overflow" in res.stderr, res.stderr[-2000:]
