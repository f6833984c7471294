"""The CPU build of the DCS physics and the oracle under the host
AddressSanitizer and UndefinedBehaviorSanitizer (g++ -fsanitize=address,
undefined, no recovery): tests/native/sanitize_host.cc drives the error
returns, serial vs OpenMP bit-identity and agreement with the oracle on
ordinary and special inputs (NaN, infinities, zero, negative, subnormal and
extreme K and q)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_host_build_and_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "sanitize_host"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fopenmp", "-ffp-contract=off", "-Wno-unknown-pragmas",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "noa_amd", "csrc"),
           os.path.join(ROOT, "tests", "native", "sanitize_host.cc"),
           os.path.join(ROOT, "noa_amd", "csrc", "dcs_host.cc"),
           os.path.join(ROOT, "oracle", "dcs_oracle.cc"), "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    run = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    print(run.stdout[-2000:])
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "ERROR: AddressSanitizer" not in run.stderr and "runtime error" not in run.stderr, run.stderr[-4000:]
    assert run.stdout.strip().endswith("ok")
