"""Top-level driver and snapshots on the GPU.

Resume property: "run 2 root steps" is bitwise equal to "run 1, snapshot,
reload into fresh arrays and globals, run 1", in the default (mixed) and
the FP64-only mode; the driver equals hold(); the FP64-only integrator is
bit-identical with the FP64 CPU HOLD."""
import numpy as np
import pytest

from conftest import plummer

pytestmark = pytest.mark.gpu


def _ic(hhs, n=300, seed=61, eps=1e-3):
    return plummer(hhs, n, seed=seed, eps2=eps * eps)


@pytest.mark.parametrize("modes", [dict(), dict(timestep_mode=1),
                                   dict(kick_precision=3)])
def test_resume_is_bitwise(hhs, torch_cuda, tmp_path, modes):
    pos, vel, m = _ic(hhs)
    dtr = 1.0 / 32
    a = hhs.HoldIntegrator(pos, vel, m, eta=0.01, eps=1e-3, dt_root=dtr, **modes)
    a.evolve_model(2 * dtr)
    b = hhs.HoldIntegrator(pos, vel, m, eta=0.01, eps=1e-3, dt_root=dtr, **modes)
    snap = str(tmp_path / "b.snap")
    b.evolve_model(dtr, snapshot_path=snap)
    b.close()
    hdr = hhs.read_snapshot_header(snap)
    assert hdr["root_steps"] == 1 and hdr["n"] == len(m) and hdr["dt_root"] == dtr
    c = hhs.HoldIntegrator.load(snap)
    assert c.model_time == dtr
    c.evolve_model(2 * dtr)
    sa, sc = a.get_state(), c.get_state()
    for x, y in zip(sa, sc):
        assert x.tobytes() == y.tobytes()
    assert a.globals.simtime == c.globals.simtime == 2 * dtr
    assert a.globals.root_steps == c.globals.root_steps == 2


def test_periodic_snapshots_hold_the_state(hhs, torch_cuda, tmp_path):
    pos, vel, m = _ic(hhs, n=200, seed=62)
    dtr = 1.0 / 64
    it = hhs.HoldIntegrator(pos, vel, m, eta=0.02, eps=1e-3, dt_root=dtr)
    snap = str(tmp_path / "p.snap")
    it.evolve_model(5 * dtr, snapshot_path=snap, snapshot_every=2)
    assert hhs.read_snapshot_header(snap)["root_steps"] == 5  # the final one
    again = hhs.HoldIntegrator.load(snap)
    for x, y in zip(it.get_state(), again.get_state()):
        assert x.tobytes() == y.tobytes()
    # a finished run continues only when asked for more time
    again.evolve_model(5 * dtr)
    assert again.globals.root_steps == 5


def test_evolve_equals_hold(hhs, torch_cuda):
    pos, vel, m = _ic(hhs, n=256, seed=63)
    d1 = hhs.DeviceParticles.from_host(pos, vel, m)
    d2 = hhs.DeviceParticles.from_host(pos, vel, m)
    g1, g2 = hhs.Globals(256, 0.01, 1e-3), hhs.Globals(256, 0.01, 1e-3)
    g1.hold(d1, 0.125, 1.0 / 32)
    g2.evolve(d2, 0.125, 1.0 / 32)
    torch_cuda.cuda.synchronize()
    assert d1.pos.cpu().numpy().tobytes() == d2.pos.cpu().numpy().tobytes()
    assert d1.vel.cpu().numpy().tobytes() == d2.vel.cpu().numpy().tobytes()
    assert g2.root_steps == 4 and g1.stats()["calls"] == g2.stats()["calls"]


def test_fp64_integrator_matches_cpu_hold_bitwise(hhs, oracle, torch_cuda):
    pos, vel, m = _ic(hhs, n=128, seed=64, eps=1e-2)
    host = oracle.HostParticles(len(m), pos, vel, m)
    rc, ost = oracle.hold(host, 0.125, 1.0 / 32, 0.01, 1e-4)
    assert rc == 0
    it = hhs.HoldIntegrator(pos, vel, m, eta=0.01, eps=1e-2, dt_root=1.0 / 32,
                            kick_precision=hhs.KICK_F64_EXACT)
    it.evolve_model(0.125)
    p, v, mm, ids = it.get_state()
    assert p.tobytes() == host.pos.tobytes() and v.tobytes() == host.vel.tobytes()
    assert np.array_equal(ids, np.arange(len(m)))
    ke, pe = it.energy()
    e = oracle.energy(host.pos, host.vel, host.m, 1e-4)
    assert abs(ke + pe - e) <= 1e-12 * abs(e)


def test_load_refuses_other_parameters_and_damage(hhs, torch_cuda, tmp_path):
    pos, vel, m = _ic(hhs, n=64, seed=65)
    it = hhs.HoldIntegrator(pos, vel, m, eta=0.01, eps=1e-3, dt_root=1.0 / 16)
    snap = tmp_path / "x.snap"
    it.evolve_model(1.0 / 16, snapshot_path=str(snap))
    d = hhs.DeviceParticles.empty(64)
    g = hhs.Globals(64, 0.02, 1e-3)  # other eta
    with pytest.raises(hhs.HHSError) as ei:
        g.load_snapshot(str(snap), d)
    assert ei.value.code == hhs.HHS_ERR_INVALID
    small = hhs.DeviceParticles.empty(63)
    with pytest.raises(hhs.HHSError) as ei:
        hhs.Globals(64, 0.01, 1e-3).load_snapshot(str(snap), small)
    assert ei.value.code == hhs.HHS_ERR_INVALID
    blob = bytearray(snap.read_bytes())
    blob[200] ^= 1  # flip one payload bit
    bad = tmp_path / "bad.snap"
    bad.write_bytes(bytes(blob))
    with pytest.raises(hhs.HHSError) as ei:
        hhs.Globals(64, 0.01, 1e-3).load_snapshot(str(bad), d)
    assert ei.value.code == hhs.HHS_ERR_FORMAT
