"""Work sharing in the ensemble kernel (parallel-hhs_amd/csrc/ensemble.hip):
workgroups that find the work queue empty run 256-sink chunks of the pair
passes of systems still running on whole workgroups.

* A batch of a few large systems (whole-workgroup teams, calls of up to
  2048 particles) among many small ones integrates to the same bits with
  sharing off, at its default threshold and at the lowest one, in the FP32
  and the exact timestep modes; the counters show that helpers ran chunks.
* The bounded wait: with a 1-microsecond bound on an owner's wait for its
  helpers, owners give up (those systems end with HHS_ERR_TIMEOUT), the
  launch still ends, and every system that finished is bit-identical to
  the run without sharing.
The protocol itself is modelled on the CPU in tests/test_share_protocol.py.

Skipped: the first GPU run of the kernel built with work sharing did not
finish within the 120 s test limit, in the first case below with sharing
switched off (share_min_n = -1) on this 43-system batch.  The cause was not
found by reading the code, so the device side was taken out again and the
library has no share_min_n / share_wait_us; these tests stay as the record
of the case and of what a next attempt has to pass.
"""
import numpy as np
import pytest

from conftest import plummer

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="ensemble work sharing removed after its first GPU run "
                                      "did not finish (cause not found); no device code to test")]

EPS, ETA = 1e-4, 0.01
T_END, DT_ROOT = 0.125, 1.0 / 16


def _systems(hhs):
    big = [2048, 1500, 1800]
    small = [8 + (k * 37) % 113 for k in range(40)]
    sizes = big + small
    return [plummer(hhs, n, seed=8100 + k, eps2=EPS * EPS) for k, n in enumerate(sizes)], sizes


def _run(hhs, torch_cuda, systems, mode, share_min_n, share_wait_us=0):
    ens = hhs.Ensemble(systems, T_END, DT_ROOT, ETA, EPS, fast_timesteps=mode,
                       share_min_n=share_min_n, share_wait_us=share_wait_us)
    print(f"ensemble: {len(systems)} systems, mode {mode}, share_min_n {share_min_n}, "
          f"wait {share_wait_us} us", flush=True)
    ens.run()
    torch_cuda.cuda.synchronize()
    return ens


def _same(a, b, k):
    ra, rb = a.results()[k], b.results()[k]
    for key in ("calls", "interactions", "ts_pairs", "max_level", "e0", "e1", "t"):
        assert ra[key] == rb[key], (k, key)
    pa, va = a.state(k)
    pb, vb = b.state(k)
    assert pa.tobytes() == pb.tobytes() and va.tobytes() == vb.tobytes(), k


@pytest.mark.parametrize("mode", ["fp32", "exact"])
def test_sharing_gives_the_same_bits(hhs, torch_cuda, mode):
    ts = hhs.TIMESTEP_FP32 if mode == "fp32" else hhs.TIMESTEP_EXACT
    systems, sizes = _systems(hhs)
    off = _run(hhs, torch_cuda, systems, ts, -1)
    assert off.sharing() == (0, 0)
    assert (off.results()["status"] == 0).all()
    for share_min in (0, 256):
        on = _run(hhs, torch_cuda, systems, ts, share_min)
        offered, helped = on.sharing()
        print(f"share_min_n={share_min}: {offered} passes offered, {helped} chunks helped")
        assert offered > 0 and helped > 0
        assert (on.results()["status"] == 0).all()
        for k in range(len(sizes)):
            _same(off, on, k)


def test_sharing_wait_bound_ends_the_system_not_the_launch(hhs, torch_cuda):
    systems, sizes = _systems(hhs)
    off = _run(hhs, torch_cuda, systems, hhs.TIMESTEP_FP32, -1)
    short = _run(hhs, torch_cuda, systems, hhs.TIMESTEP_FP32, 256, share_wait_us=1)
    st = short.results()["status"]
    assert set(np.unique(st).tolist()) <= {0, hhs.HHS_ERR_TIMEOUT}
    assert (st == hhs.HHS_ERR_TIMEOUT).any(), "no owner ran out of its 1 us bound"
    # only whole-workgroup systems share, so only they can time out
    assert all(sizes[k] > 512 for k in np.nonzero(st)[0])
    for k in np.nonzero(st == 0)[0]:
        _same(off, short, int(k))
