"""The HIP path against the golden fixtures (tests/golden/hold_golden.npz,
FP64 oracle outputs on seeded inputs): the F64_EXACT kick, EXACT
timesteps, the split, the scan and the FP64-only HOLD step bit for bit;
the MIXED32 kick within 1e-6 of sum_j m_j/(r_ij^2 + eps^2) per sink and
FP32 timesteps within 1e-6 relative.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import force_scale

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "hold_golden.npz")
CASES = ("n8", "n64", "n1024", "n64s")
EPS2S = (0.0, 1e-8)
ETA = 0.01


@pytest.fixture(scope="module")
def golden():
    with np.load(PATH, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _dev(hhs, g, name, ts=None):
    return hhs.DeviceParticles.from_host(g[f"{name}_pos"], g[f"{name}_vel"], g[f"{name}_m"], ts)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("k", [0, 1])
def test_golden_kick(hhs, torch_cuda, golden, name, k):
    eps2 = EPS2S[k]
    ref = golden[f"{name}_acc_e{k}"]
    pos, m = golden[f"{name}_pos"], golden[f"{name}_m"]
    d = _dev(hhs, golden, name)
    hhs.kick_self(d, dt=1e-9, fac=1.0, eps2=eps2)
    torch_cuda.cuda.synchronize()
    err = np.linalg.norm(d.acc.cpu().numpy() - ref, axis=1) / force_scale(pos, m, eps2)
    assert err.max() < 1e-6
    d = _dev(hhs, golden, name)
    hhs.kick_set_precision(hhs.KICK_F64_EXACT)
    try:
        hhs.kick_self(d, dt=1e-9, fac=1.0, eps2=eps2)
        torch_cuda.cuda.synchronize()
    finally:
        hhs.kick_set_precision(hhs.KICK_MIXED32)
    assert d.acc.cpu().numpy().tobytes() == ref.tobytes()


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("k", [0, 1])
def test_golden_timesteps(hhs, torch_cuda, golden, name, k):
    eps2, ref = EPS2S[k], golden[f"{name}_ts_e{k}"]
    d = _dev(hhs, golden, name)
    hhs.timesteps(d, eta=ETA, eps2=eps2)
    torch_cuda.cuda.synchronize()
    assert d.timestep.cpu().numpy().tobytes() == ref.tobytes()
    hhs.timesteps(d, eta=ETA, eps2=eps2, fast=hhs.TIMESTEP_FP32)
    torch_cuda.cuda.synchronize()
    assert np.all(np.abs(d.timestep.cpu().numpy() - ref) <= 1e-6 * ref)


@pytest.mark.parametrize("p", [3, 6, 8, 10, 12, 13, 14])
def test_golden_split(hhs, torch_cuda, golden, p):
    torch = torch_cuda
    n = len(golden["n1024_m"])
    d = _dev(hhs, golden, "n1024", golden["n1024_ts_e0"])
    slow, fast = hhs.DeviceParticles.empty(n), hhs.DeviceParticles.empty(n)
    part = torch.zeros(2, dtype=torch.int32, device="cuda")
    hhs.split(d, slow, fast, 2.0 ** -p, part, hhs.ScanWorkspace(n))
    torch.cuda.synchronize()
    ns, nf = (int(x) for x in part.cpu())
    assert [ns, nf] == golden[f"split_p{p}_counts"].tolist()
    assert np.array_equal(fast.parent_id.cpu().numpy()[:nf], golden[f"split_p{p}_fast_parent"])
    assert np.array_equal(slow.parent_id.cpu().numpy()[:ns], golden[f"split_p{p}_slow_parent"])


@pytest.mark.parametrize("n", [512, 1023, 2048, 2151, 4090, 588021])
def test_golden_scan(hhs, torch_cuda, golden, n):
    torch = torch_cuda
    x = (np.random.default_rng(1000 + n).random(n) * 10.0 <= 5.0).astype(np.int32)
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    hhs.scan_exclusive(torch.as_tensor(x).cuda(), out, hhs.ScanWorkspace(n))
    torch.cuda.synchronize()
    s = out.cpu().numpy()
    assert int(s[-1]) + int(x[-1]) == int(golden[f"scan_{n}_total"][0])
    assert hashlib.sha256(s.tobytes()).digest() == golden[f"scan_{n}_sha256"].tobytes()


@pytest.mark.parametrize("name", ["n64", "n1024"])
def test_golden_hold_step(hhs, torch_cuda, golden, name):
    dt = 1.0 / 32
    n = len(golden[f"{name}_m"])
    d = _dev(hhs, golden, name)
    g = hhs.Globals(n, ETA, 0.0, kick_precision=hhs.KICK_F64_EXACT)
    g.hold_step(d, dt)
    torch_cuda.cuda.synchronize()
    assert g.stats()["calls_per_level"] == golden[f"hold_{name}_calls"].tolist()
    assert g.simtime == golden[f"hold_{name}_simtime"][0]
    for f, key in (("pos", "pos"), ("vel", "vel"), ("timestep", "ts")):
        assert getattr(d, f).cpu().numpy().tobytes() == golden[f"hold_{name}_{key}"].tobytes()
    g.close()
