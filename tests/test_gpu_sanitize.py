"""The server's host code (scheduler thread, workers, host-mapped hand-back,
per-slot buffers) under AddressSanitizer and UndefinedBehaviorSanitizer:
tests/cpp/server_asan.cpp, linked against lib/libhhs_asan.so (`make -C
parallel-hhs_amd asan`, which __graft_entry__.build() runs; device code is
not instrumented), submits from two threads, waits, fetches, releases,
refuses an oversized request and destroys a server with requests in flight.
It must exit 0 with no sanitizer report.  Leak reports from the
uninstrumented ROCm runtime libraries are suppressed
(tests/cpp/lsan_runtime.supp); leaks in libhhs are not."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "server_asan")
SUPP = os.path.join(ROOT, "tests", "cpp", "lsan_runtime.supp")

pytestmark = pytest.mark.gpu


def test_server_host_code_clean_under_asan_ubsan():
    assert os.path.exists(EXE), "run `make -C parallel-hhs_amd asan` (or build()) first"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               LSAN_OPTIONS=f"suppressions={SUPP}", UBSAN_OPTIONS="halt_on_error=1")
    print("server_asan: start", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:], flush=True)
    assert r.returncode == 0
    assert "server sanitizer driver: ok" in r.stdout
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
