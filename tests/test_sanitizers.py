"""Sanitizer builds of the host code and the oracle (CPU only; gcc/g++).

  * the drop-in driver's parsing, output writers and -s / -l files, and the serving
    scheduler, built with g++ over a host-only stand-in for the engine ABI
    (tests/sanitize/fake_engine.cpp), under AddressSanitizer + UndefinedBehaviorSanitizer
    and under ThreadSanitizer;
  * the CPU oracle's threaded paths (every sampling mode, partner backgrounds, the
    long-horizon G(t) runner, the timing runner) under ASan + UBSan and under TSan.
Each run must exit 0 with no sanitizer report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
CSRC = os.path.join(ROOT, "gpu_dsm_amd", "csrc")
ASAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined"]
TSAN = ["-fsanitize=thread"]


def _build_run(cmd, exe, args, tmp_path):
    r = subprocess.run(cmd + ["-o", exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1",
               TSAN_OPTIONS="halt_on_error=1")
    run = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300, env=env, cwd=str(tmp_path))
    out = run.stdout + run.stderr
    assert run.returncode == 0, out[-4000:]
    for bad in ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer", "ERROR: LeakSanitizer"):
        assert bad not in out, out[-4000:]
    return out


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
@pytest.mark.parametrize("flags", [ASAN, TSAN], ids=["asan_ubsan", "tsan"])
def test_host_driver_and_serving_sanitized(tmp_path, flags):
    exe = str(tmp_path / "host_checks")
    cmd = ["g++", "-std=c++17", "-g", "-O1", *flags, "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "dsm_driver.cpp"), os.path.join(CSRC, "dsm_serve.cpp"),
           os.path.join(SAN, "fake_engine.cpp"), os.path.join(SAN, "host_checks.cpp")]
    work = tmp_path / "work"
    work.mkdir()
    out = _build_run(cmd, exe, [str(work)], tmp_path)
    assert "host checks: 0 failures" in out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
@pytest.mark.parametrize("flags", [ASAN, TSAN], ids=["asan_ubsan", "tsan"])
def test_oracle_sanitized(tmp_path, flags):
    exe = str(tmp_path / "oracle_checks")
    cmd = ["gcc", "-std=gnu11", "-g", "-O1", "-ffp-contract=off", *flags, "-I", os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "oracle", "dsm_oracle.c"), os.path.join(SAN, "oracle_checks.c"), "-lm", "-lpthread"]
    out = _build_run(cmd, exe, [], tmp_path)
    assert "oracle checks: 0 failures" in out
