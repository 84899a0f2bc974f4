"""Batched (ensemble) HOLD: many systems in one persistent launch.

Each system is checked against the FP64 CPU HOLD oracle with the same bar as
the per-system path (identical recursion, positions and velocities within
1e-6 of the rms); a system must integrate to the same bits alone and inside
a mixed batch; the level guard must stop one bad system without disturbing
the others.
"""
import numpy as np
import pytest

from conftest import plummer

pytestmark = pytest.mark.gpu


def _systems(hhs, sizes, seed0, eps):
    return [plummer(hhs, n, seed=seed0 + k, eps2=eps * eps) for k, n in enumerate(sizes)]


def test_ensemble_matches_oracle_per_system(hhs, oracle, torch_cuda):
    eps, eta, t_end, dtr = 1e-2, 0.01, 1.0 / 16, 1.0 / 32
    sizes = [2, 16, 64, 100, 257, 300, 700]
    systems = _systems(hhs, sizes, 500, eps)
    ens = hhs.Ensemble(systems, t_end, dtr, eta, eps)
    ens.run()
    torch_cuda.cuda.synchronize()
    res = ens.results()
    for k, (p0, v0, m) in enumerate(systems):
        host = oracle.HostParticles(len(m), p0, v0, m)
        rc, ost = oracle.hold(host, t_end, dtr, eta, eps * eps)
        assert rc == 0
        r = res[k]
        assert r["status"] == 0
        assert r["calls"] == ost.calls, f"system {k} (n={len(m)})"
        assert r["max_level"] == ost.max_level
        assert r["t"] == pytest.approx(t_end, rel=1e-15)
        p, v = ens.state(k)
        assert np.abs(p - host.pos).max() <= 1e-6 * np.sqrt((host.pos ** 2).mean()), k
        assert np.abs(v - host.vel).max() <= 1e-6 * np.sqrt((host.vel ** 2).mean()), k
        assert r["e0"] == pytest.approx(oracle.energy(p0, v0, m, eps * eps), rel=1e-12)
        assert r["e1"] == pytest.approx(oracle.energy(p, v, m, eps * eps), rel=1e-12)


def test_ensemble_matches_per_system_gpu_path(hhs, torch_cuda):
    eps, eta, t_end, dtr = 1e-3, 0.01, 0.125, 0.125
    sizes = [40, 500]
    systems = _systems(hhs, sizes, 900, eps)
    ens = hhs.Ensemble(systems, t_end, dtr, eta, eps)
    ens.run()
    torch_cuda.cuda.synchronize()
    res = ens.results()
    for k, (p0, v0, m) in enumerate(systems):
        d = hhs.DeviceParticles.from_host(p0, v0, m)
        g = hhs.Globals(len(m), eta, eps)
        g.hold(d, t_end, dtr)
        torch_cuda.cuda.synchronize()
        assert res[k]["calls"] == g.stats()["calls"]
        p, v = ens.state(k)
        pr = d.pos.cpu().numpy()
        assert np.abs(p - pr).max() <= 1e-6 * np.sqrt((pr ** 2).mean())


def test_ensemble_alone_equals_in_mix(hhs, torch_cuda):
    """Continuous-batching invariant: a system's result is bit-identical
    whether it runs alone or inside a mixed batch."""
    eps, eta = 1e-4, 0.01
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(8), np.log(2048), 48)).astype(int)
    systems = _systems(hhs, sizes, 70, eps)
    t_end = rng.uniform(0.05, 0.15, len(sizes))
    mix = hhs.Ensemble(systems, t_end, 1.0 / 64, eta, eps)
    mix.run()
    torch_cuda.cuda.synchronize()
    rmix = mix.results()
    for k in (0, 7, int(np.argmax(sizes)), int(np.argmin(sizes))):
        alone = hhs.Ensemble([systems[k]], t_end[k], 1.0 / 64, eta, eps)
        alone.run()
        torch_cuda.cuda.synchronize()
        pa, va = alone.state(0)
        pm, vm = mix.state(k)
        assert np.array_equal(pa, pm) and np.array_equal(va, vm), f"system {k}"
        ra = alone.results()[0]
        assert ra.tobytes() == rmix[k].tobytes()


def test_ensemble_depth_guard_isolated(hhs, torch_cuda):
    eps, eta = 0.0, 0.01
    systems = _systems(hhs, [16, 16, 300], 33, 0.0)
    p, v, m = systems[1]
    p[5] = p[4]  # coincident pair, eps = 0: tau = 0, the recursion cannot end
    ens = hhs.Ensemble(systems, 1.0 / 64, 1.0 / 64, eta, eps, max_level=12)
    ens.run()
    torch_cuda.cuda.synchronize()
    res = ens.results()
    assert res[1]["status"] == hhs.HHS_ERR_DEPTH
    assert res[0]["status"] == 0 and res[2]["status"] == 0


def test_ensemble_tiny_systems_and_momentum(hhs, torch_cuda):
    eps, eta = 1e-3, 0.01
    systems = [(np.zeros((1, 3)), np.array([[0.1, 0.0, 0.0]]), np.ones(1))]
    systems += _systems(hhs, [3, 1024], 11, eps)
    ens = hhs.Ensemble(systems, 0.25, 0.125, eta, eps)
    ens.run()
    torch_cuda.cuda.synchronize()
    res = ens.results()
    assert np.all(res["status"] == 0)
    p, v = ens.state(0)
    assert np.allclose(p, [[0.025, 0.0, 0.0]], rtol=0, atol=1e-15)  # free particle drifts
    for k in (1, 2):
        p, v = ens.state(k)
        m = systems[k][2]
        mom = (m[:, None] * v).sum(0)
        assert np.abs(mom).max() <= 1e-6 * (m[:, None] * np.abs(v)).sum(0).max()


@pytest.mark.parametrize("mode", [1, 2])
def test_ensemble_timestep_modes_track_oracle(hhs, oracle, torch_cuda, mode):
    eps, eta, t_end, dtr = 1e-2, 0.01, 1.0 / 16, 1.0 / 32
    systems = _systems(hhs, [33, 300], 610, eps)
    ens = hhs.Ensemble(systems, t_end, dtr, eta, eps, fast_timesteps=mode)
    ens.run()
    torch_cuda.cuda.synchronize()
    res = ens.results()
    for k, (p0, v0, m) in enumerate(systems):
        host = oracle.HostParticles(len(m), p0, v0, m)
        oracle.hold(host, t_end, dtr, eta, eps * eps)
        assert res[k]["status"] == 0
        p, _ = ens.state(k)
        assert np.abs(p - host.pos).max() <= 1e-6 * np.sqrt((host.pos ** 2).mean())


def test_ensemble_team_size_does_not_change_bits(hhs, torch_cuda):
    """The same systems on one-wave teams (wave_team_max_n large) and on
    256-thread workgroup teams (wave_team_max_n small): identical positions,
    velocities, energies and statistics."""
    eps, eta = 1e-4, 0.01
    systems = _systems(hhs, [70, 300, 700, 1100], 1300, eps)
    runs = []
    for wave_max in (8, 4096):
        ens = hhs.Ensemble(systems, 0.125, 1.0 / 32, eta, eps, wave_team_max_n=wave_max,
                           fast_timesteps=hhs.TIMESTEP_FP32)
        ens.run()
        torch_cuda.cuda.synchronize()
        runs.append((ens.results(), [ens.state(k) for k in range(len(systems))]))
    (ra, sa), (rb, sb) = runs
    assert ra.tobytes() == rb.tobytes()
    for (pa, va), (pb, vb) in zip(sa, sb):
        assert np.array_equal(pa, pb) and np.array_equal(va, vb)


def test_ensemble_same_bits_on_new_streams(hhs, torch_cuda):
    """The first launch of each kernel variant on a stream runs a warm-up
    dispatch (no items) before the batch; the batch's bits must not depend
    on it: the same batch on the current stream and on fresh streams, in
    the FP32 and the exact timestep modes."""
    eps, eta, t_end, dtr = 1e-3, 0.01, 0.125, 1.0 / 16
    sizes = [9, 40, 130, 513, 700]
    systems = _systems(hhs, sizes, 1300, eps)
    for mode in (hhs.TIMESTEP_FP32, hhs.TIMESTEP_EXACT):
        runs = []
        for stream in (None, torch_cuda.cuda.Stream(), torch_cuda.cuda.Stream(), None):
            ens = hhs.Ensemble(systems, t_end, dtr, eta, eps, fast_timesteps=mode)
            ens.run(stream)
            torch_cuda.cuda.synchronize()
            assert (ens.results()["status"] == 0).all()
            runs.append([ens.state(k) for k in range(len(sizes))])
        for other in runs[1:]:
            for (pa, va), (pb, vb) in zip(runs[0], other):
                assert pa.tobytes() == pb.tobytes() and va.tobytes() == vb.tobytes()
