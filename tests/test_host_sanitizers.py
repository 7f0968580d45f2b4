"""Host AddressSanitizer + UndefinedBehaviorSanitizer runs (CPU only).

* The oracle (oracle/orc.c) built with -fsanitize=address,undefined and driven
  by oracle/sanitize_main.c over every routine on the icosahedron with padded
  leading dimensions, wave-number chunks and a row subset.
* The library's host code (C-ABI argument checks, mesh validation, host
  geometry, parameter loading, request estimates) built with each
  -fsanitize= applied to the host pass only and driven by
  baemm_amd/csrc/sanitize_host.cpp, which needs no GPU.

A sanitizer finding aborts the driver with a report; the drivers' own checks
print FAILED lines."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    print(r.stdout[-4000:], r.stderr[-4000:])
    return r


def test_oracle_under_asan_and_ubsan():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    assert _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"]).returncode == 0
    r = _run([os.path.join(ROOT, "oracle", "_asan", "orc_asan")], timeout=300)
    assert r.returncode == 0 and "OK" in r.stdout


def test_capi_host_paths_under_asan_and_ubsan():
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    assert _run(["make", "-s", "-j4", "-C", os.path.join(ROOT, "baemm_amd", "csrc"), "asan"],
                timeout=1200).returncode == 0
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = _run([os.path.join(ROOT, "_build", "asan", "sanitize_host")], timeout=300, env=env)
    assert r.returncode == 0 and "0 failed checks" in r.stdout
