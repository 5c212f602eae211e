"""Host-side sanitizer runs (CPU only, no GPU needed):

* the oracle's C restatements (complex and real) built with AddressSanitizer
  and UndefinedBehaviorSanitizer (``make -C oracle sanitize``) and driven over
  edge sizes, non-finite input, a zero sweep cap, balancing jobs and the
  Schur / adjoint / eigenvector entries (oracle/sanitize_driver.c);
* the library's C++ host layer: every library source compiled with ASan and
  UBSan on the host side only (``-Xarch_host``: device code is not
  instrumented) and linked with tests/host_sanitize_driver.cpp, which runs
  handle creation/teardown, argument validation of every entry point,
  workspace sizing, the ragged schedule builder on valid and malformed
  offset arrays, and the tuning-cache parser.
A sanitizer report aborts the driver, so a nonzero exit fails the test.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")


def test_oracle_restatements_under_asan_ubsan():
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    r = subprocess.run([os.path.join(ROOT, "oracle", "sanitize_driver")], capture_output=True, text=True, env=ENV,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "no report" in r.stdout


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_host_layer_under_asan_ubsan(tmp_path):
    import tokura_zgeev_batched_amd.build as b
    csrc = os.path.join(ROOT, "tokura_zgeev_batched_amd", "csrc")
    flags = ["-O1", "-g", "-std=c++17", "-fPIC", f"--offload-arch={b.ARCH}", "-I", os.path.join(ROOT, "include"),
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]

    def compile_one(src):
        obj = str(tmp_path / (os.path.splitext(os.path.basename(src))[0] + ".o"))
        r = subprocess.run([HIPCC, *flags, *SAN, "-x", "hip", "-c", src, "-o", obj], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]
        return obj

    srcs = [os.path.join(csrc, s) for s in b.SOURCES] + [os.path.join(ROOT, "tests", "host_sanitize_driver.cpp")]
    with ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        objs = list(ex.map(compile_one, srcs))
    exe = str(tmp_path / "host_sanitize_driver")
    r = subprocess.run([HIPCC, f"--offload-arch={b.ARCH}", *SAN[:4], "-o", exe, *objs], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe, str(tmp_path / "tune.txt")], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "no report" in r.stdout
    shutil.rmtree(tmp_path, ignore_errors=True)
