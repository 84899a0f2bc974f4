"""Golden fixtures of the HOLD path (tests/golden/hold_golden.npz, written by
tests/golden/make_golden.py from the FP64 oracle): the oracle and the
library's initial-condition generator still reproduce every array bit for
bit, and the fixtures satisfy the invariants the mathematics fixes.
"""
import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "hold_golden.npz")


def _maker():
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def golden():
    with np.load(PATH, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_golden_reproduced_bit_for_bit(hhs, oracle, golden):
    new = _maker().generate(hhs, oracle)
    assert sorted(new) == sorted(golden)
    for k, ref in golden.items():
        got = np.asarray(new[k])
        assert got.dtype == ref.dtype and got.shape == ref.shape, k
        assert got.tobytes() == ref.tobytes(), k


def test_golden_invariants(golden):
    mk = _maker()
    for name, n, _, mass_model in mk.IC_CASES:
        m = golden[f"{name}_m"]
        assert m.shape == (n,) and abs(m.sum() - 1.0) < 1e-12
        if mass_model == 0:
            assert np.all(m == 1.0 / n)
        for k in range(len(mk.EPS2S)):
            acc = golden[f"{name}_acc_e{k}"]
            # pairwise forces are antisymmetric: sum m_i a_i = 0 to roundoff
            scale = (m[:, None] * np.abs(acc)).sum(0).max()
            assert np.abs((m[:, None] * acc).sum(0)).max() <= 1e-12 * scale
            assert np.all(golden[f"{name}_ts_e{k}"] > 0)
    ts = golden["n1024_ts_e0"]
    for p in mk.PIVOT_EXPS:
        ns, nf = (int(x) for x in golden[f"split_p{p}_counts"])
        fast, slow = golden[f"split_p{p}_fast_parent"], golden[f"split_p{p}_slow_parent"]
        raw = int((ts <= 2.0 ** -p).sum())
        assert nf == (0 if raw == 1 else raw) and ns + nf == len(ts)
        # stable partition: both parent lists ascending, together a permutation
        assert np.all(np.diff(fast) > 0) and np.all(np.diff(slow) > 0)
        assert np.array_equal(np.sort(np.concatenate([fast, slow])), np.arange(len(ts)))
    for name in mk.HOLD_CASES:
        calls = golden[f"hold_{name}_calls"]
        assert calls[0] == 1 and np.all(calls[1:] <= 2 * calls[:-1])
        assert golden[f"hold_{name}_simtime"][0] == mk.HOLD_DT
        m, v0, v1 = golden[f"{name}_m"], golden[f"{name}_vel"], golden[f"hold_{name}_vel"]
        # momentum is conserved by the HOLD step (S<-F and F<-S kicks pair up)
        assert np.abs((m[:, None] * (v1 - v0)).sum(0)).max() <= 1e-13
    for n in mk.SCAN_SIZES:
        assert golden[f"scan_{n}_total"][0] == int(mk.scan_input(n).sum())
