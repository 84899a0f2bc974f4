"""Host sanitizer build of the FP64 oracle: tests/cpp/oracle_sanitize.c drives
every oracle entry point on small inputs and edge cases, compiled together
with oracle/hhs_oracle.c under AddressSanitizer and UndefinedBehaviorSanitizer
(gcc).  It must run clean, and a planted write past an array must be caught.
The server's host code runs under the same sanitizers on the GPU
(tests/test_gpu_sanitize.py)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "cpp", "oracle_sanitize.c"),
       os.path.join(ROOT, "oracle", "hhs_oracle.c")]
FLAGS = ["-std=c11", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-fno-omit-frame-pointer", "-ffp-contract=off", "-fopenmp",
         "-I" + os.path.join(ROOT, "oracle")]

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")


def _build_run(tmp_path, *defs):
    exe = str(tmp_path / "osan")
    subprocess.run(["gcc", *FLAGS, *defs, *SRC, "-lm", "-o", exe], check=True,
                   capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", OMP_NUM_THREADS="2")
    return subprocess.run([exe], capture_output=True, text=True, timeout=120, env=env)


def test_oracle_clean_under_asan_ubsan(tmp_path):
    r = _build_run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "oracle sanitizer driver: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_planted_overflow_is_caught(tmp_path):
    r = _build_run(tmp_path, "-DPLANT_OOB")
    assert r.returncode != 0
    assert "AddressSanitizer" in r.stderr
