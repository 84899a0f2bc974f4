"""Host-thread model of a work-sharing protocol for the ensemble kernel
(tests/cpp/share_model.cpp; its device side was taken out after its first
GPU run did not finish, tests/test_gpu_ensemble_sharing.py).  CPU only: the
model is built with g++, plain and under ThreadSanitizer, and must

* give every pass's outputs exactly (each chunk run once, by the owner or a
  helper, with the descriptor of the pass it claimed) with no data race;
* end when a helper stalls past the owner's bound, with exactly that pass
  reported failed;
* fail when a mistake is planted in it: the owner not waiting for the done
  count, a done increment without release, a claim word without its
  sequence number.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "share_model.cpp")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _build(tmp_path, name, *flags):
    out = str(tmp_path / name)
    subprocess.run(["g++", "-std=c++17", "-pthread", "-g", *flags, SRC, "-o", out], check=True,
                   capture_output=True, text=True)
    return out


def _run(exe, *args, timeout=120):
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


@pytest.fixture(scope="module")
def binaries(tmp_path_factory):
    d = tmp_path_factory.mktemp("share_model")
    return {"plain": _build(d, "plain", "-O2"),
            "tsan": _build(d, "tsan", "-O1", "-fsanitize=thread")}


@pytest.mark.parametrize("owners,helpers,jobs,seed,fixed_n",
                         [(6, 6, 300, 1, 0), (2, 10, 300, 2, 0), (8, 2, 200, 3, 0),
                          (4, 4, 300, 4, 600)])
def test_protocol_outputs_exact(binaries, owners, helpers, jobs, seed, fixed_n):
    r = _run(binaries["plain"], owners, helpers, jobs, seed, 0, fixed_n)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " bad 0 " in r.stdout and " failed 0 " in r.stdout
    assert "helper_claims 0 " not in r.stdout  # the helpers did take work


@pytest.mark.parametrize("seed,fixed_n", [(11, 0), (12, 600)])
def test_protocol_race_free_under_tsan(binaries, seed, fixed_n):
    r = _run(binaries["tsan"], 4, 4, 150, seed, 0, fixed_n)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ThreadSanitizer" not in r.stderr


def test_protocol_ends_when_a_helper_stalls(binaries):
    """One helper stalls after claiming for longer than the owner's bound:
    the owner gives up on that pass (reported failed, slot retired) and the
    run ends; every other pass is exact."""
    r = _run(binaries["plain"], 4, 4, 40, 21, 1)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " failed 1 " in r.stdout and " bad 0 " in r.stdout


@pytest.mark.parametrize("plant,fixed_n", [("PLANT_NO_WAIT", 0), ("PLANT_RELAXED_DONE", 0),
                                           ("PLANT_NO_SEQ", 600)])
def test_planted_mistakes_are_caught(tmp_path, plant, fixed_n):
    exe = _build(tmp_path, plant, "-O1", "-fsanitize=thread", f"-D{plant}")
    r = _run(exe, 4, 4, 100, 5, 0, fixed_n)
    assert r.returncode != 0, f"{plant} went unnoticed: {r.stdout}"
