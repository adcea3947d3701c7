"""Host code of the C ABI (the implicit-QR tridiagonal eigensolver that stays on the CPU, its range
forms and the complex-tridiagonal reduction) built with AddressSanitizer and
UndefinedBehaviorSanitizer together with the oracle, and driven by tests/sanitize/host_sanitize.cpp:
sizes 0..129, split matrices, the stagnation fallback and argument errors, with the results compared
to the oracle bit for bit (or by the Frobenius invariant where the oracle would call LAPACK)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="no host compiler")
def test_host_code_under_asan_ubsan(tmp_path):
    obj = str(tmp_path / "oracle.o")
    exe = str(tmp_path / "host_sanitize")
    subprocess.run(["gcc", "-c", "-std=c11", "-ffp-contract=off"] + SAN +
                   [os.path.join(ROOT, "oracle", "skew_oracle.c"), "-o", obj], check=True)
    subprocess.run(["g++", "-std=c++17", "-Wno-unknown-pragmas"] + SAN +
                   [os.path.join(HERE, "sanitize", "host_sanitize.cpp"),
                    os.path.join(ROOT, "skewlinearalgebra.jl_amd", "csrc", "host_tridiag.cpp"), obj,
                    "-o", exe, "-lm", "-ldl"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "clean (0 failures)" in r.stdout
