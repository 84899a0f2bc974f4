"""HOLD recursion on the GPU against the FP64 CPU HOLD oracle.

Tolerances: positions and velocities after one root step agree to 1e-6
relative to the system's rms values (FP32 forces vs FP64); the recursion
(calls per level, simulated time) is identical; the relative energy error
|dE/E| after a fixed integration time agrees with the FP64 HOLD run to
within 10%, measured at eta = 0.1 where the integrator error dominates the
FP32 force rounding.
"""
import numpy as np
import pytest

from conftest import plummer

pytestmark = pytest.mark.gpu


def _run_pair(hhs, oracle, pos, vel, m, eta, eps, t_end, dt_root, subtree_max=None):
    eps2 = eps * eps
    host = oracle.HostParticles(len(m), pos, vel, m)
    rc, ost = oracle.hold(host, t_end, dt_root, eta, eps2)
    assert rc == 0
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(len(m), eta, eps, subtree_max=subtree_max)
    g.hold(d, t_end, dt_root)
    return host, ost, d, g


# subtree_max 0: host recursion for every call; 64: calls with <= 64
# particles run their subtree on the device; 4096: the whole root step does.
@pytest.mark.parametrize("subtree_max", [0, 64, 4096])
@pytest.mark.parametrize("n", [2, 16, 128, 512])
def test_hold_one_root_step_matches_oracle(hhs, oracle, torch_cuda, n, subtree_max):
    eps, eta = 1e-2, 0.01
    pos, vel, m = plummer(hhs, n, seed=1000 + n, eps2=eps * eps)
    host, ost, d, g = _run_pair(hhs, oracle, pos, vel, m, eta, eps, 1.0 / 32, 1.0 / 32,
                                subtree_max)
    torch_cuda.cuda.synchronize()
    st = g.stats()
    assert st["calls"] == ost.calls and st["interactions"] == ost.interactions
    assert st["ts_pairs"] == ost.ts_pairs and st["max_level"] == ost.max_level
    assert st["calls_per_level"] == [int(x) for x in ost.calls_per_level[:ost.max_level + 1]]
    assert g.simtime == pytest.approx(ost.simtime, rel=0, abs=0)
    p, v = d.pos.cpu().numpy(), d.vel.cpu().numpy()
    assert np.abs(p - host.pos).max() <= 1e-6 * np.sqrt((host.pos**2).mean())
    assert np.abs(v - host.vel).max() <= 1e-6 * np.sqrt((host.vel**2).mean())
    assert np.array_equal(d.m.cpu().numpy(), host.m)


@pytest.mark.parametrize("n,eps", [(2, 0.0), (16, 1e-2), (129, 1e-2), (512, 1e-3), (1024, 0.0)])
def test_hold_fp64_mode_bit_exact(hhs, oracle, torch_cuda, n, eps):
    """FP64-only HOLD (exact timesteps + F64_EXACT kicks): the whole
    integration -- recursion, positions, velocities, timesteps, simulated
    time -- is bit-identical with the FP64 CPU HOLD."""
    eta, t_end, dtr = 0.01, 1.0 / 16, 1.0 / 32
    pos, vel, m = plummer(hhs, n, seed=2000 + n, eps2=eps * eps)
    host = oracle.HostParticles(n, pos, vel, m)
    rc, ost = oracle.hold(host, t_end, dtr, eta, eps * eps)
    assert rc == 0
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(n, eta, eps, kick_precision=hhs.KICK_F64_EXACT)
    g.hold(d, t_end, dtr)
    torch_cuda.cuda.synchronize()
    st = g.stats()
    assert st["calls"] == ost.calls and st["interactions"] == ost.interactions
    assert g.simtime == ost.simtime
    assert d.pos.cpu().numpy().tobytes() == host.pos.tobytes()
    assert d.vel.cpu().numpy().tobytes() == host.vel.tobytes()
    assert d.timestep.cpu().numpy().tobytes() == host.timestep.tobytes()


def test_hold_fast_timesteps_matches_oracle(hhs, oracle, torch_cuda):
    """HOLD with FAST timesteps: same recursion and trajectories to 1e-6."""
    n, eps, eta = 256, 1e-2, 0.01
    pos, vel, m = plummer(hhs, n, seed=4242, eps2=eps * eps)
    host = oracle.HostParticles(n, pos, vel, m)
    rc, ost = oracle.hold(host, 1.0 / 16, 1.0 / 32, eta, eps * eps)
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(n, eta, eps, fast_timesteps=True)
    g.hold(d, 1.0 / 16, 1.0 / 32)
    torch_cuda.cuda.synchronize()
    assert g.stats()["calls"] == ost.calls
    p = d.pos.cpu().numpy()
    assert np.abs(p - host.pos).max() <= 1e-6 * np.sqrt((host.pos**2).mean())


def test_hold_fp32_timesteps_tracks_oracle(hhs, oracle, torch_cuda):
    """HOLD with FP32 timesteps: trajectories within 1e-6 of the rms and the
    same energy error as the FP64 HOLD run to 10% (eta = 0.1)."""
    n, eps, eta = 256, 1e-2, 0.1
    pos, vel, m = plummer(hhs, n, seed=4343, eps2=eps * eps)
    e0 = oracle.energy(pos, vel, m, eps * eps)
    host = oracle.HostParticles(n, pos, vel, m)
    rc, ost = oracle.hold(host, 0.25, 1.0 / 16, eta, eps * eps)
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(n, eta, eps, fast_timesteps=hhs.TIMESTEP_FP32)
    g.hold(d, 0.25, 1.0 / 16)
    torch_cuda.cuda.synchronize()
    p = d.pos.cpu().numpy()
    assert np.abs(p - host.pos).max() <= 1e-6 * np.sqrt((host.pos**2).mean())
    ke, pe = hhs.energy(d.pos, d.vel, d.m, eps * eps)
    de_cpu = abs(oracle.energy(host.pos, host.vel, host.m, eps * eps) / e0 - 1)
    assert abs(abs((ke + pe) / e0 - 1) - de_cpu) <= 0.1 * de_cpu


def test_hold_energy_error_matches_fp64_run(hhs, oracle, torch_cuda):
    n, eps, eta = 256, 1e-2, 0.1
    pos, vel, m = plummer(hhs, n, seed=77, eps2=eps * eps)
    e0 = oracle.energy(pos, vel, m, eps * eps)
    host, ost, d, g = _run_pair(hhs, oracle, pos, vel, m, eta, eps, 0.5, 1.0 / 16)
    e_cpu = oracle.energy(host.pos, host.vel, host.m, eps * eps)
    ke, pe = hhs.energy(d.pos, d.vel, d.m, eps * eps)
    de_cpu = abs((e_cpu - e0) / e0)
    de_gpu = abs((ke + pe - e0) / e0)
    assert de_cpu > 1e-9, "integrator error too small for a meaningful comparison"
    assert abs(de_gpu - de_cpu) <= 0.1 * de_cpu, f"|dE/E| gpu {de_gpu:.4e} cpu {de_cpu:.4e}"


def test_hold_kepler_period(hhs, torch_cuda):
    """Eccentric two-body orbit (a = 1, e = 0.5, M = 1): after one period
    T = 2*pi the bodies return to their start (second-order error)."""
    e, mtot = 0.5, 1.0
    rp = 1.0 - e
    vp = np.sqrt(mtot * (1 + e) / rp)
    m = np.array([0.5, 0.5])
    pos = np.array([[-rp / 2, 0, 0], [rp / 2, 0, 0]])
    vel = np.array([[0, -vp / 2, 0], [0, vp / 2, 0]])
    T = 2 * np.pi
    nsteps = 4096
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(2, 0.01, 0.0)
    for _ in range(nsteps):
        g.hold_step(d, T / nsteps)
    torch_cuda.cuda.synchronize()
    p = d.pos.cpu().numpy()
    assert np.abs(p - pos).max() < 1e-3


def test_hold_momentum_conserved(hhs, torch_cuda):
    n, eps = 300, 1e-3
    pos, vel, m = plummer(hhs, n, seed=8, mass_model=1, eps2=eps * eps)
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(n, 0.01, eps)
    g.hold(d, 0.125, 0.125)
    torch_cuda.cuda.synchronize()
    p1 = (m[:, None] * d.vel.cpu().numpy()).sum(0)
    scale = (m[:, None] * np.abs(d.vel.cpu().numpy())).sum(0)
    assert np.all(np.abs(p1) <= 1e-6 * scale)


@pytest.mark.parametrize("subtree_max", [0, 4096])
def test_hold_depth_guard(hhs, torch_cuda, subtree_max):
    """Two coincident particles with eps = 0 give tau = 0: the reference
    would recurse forever; here the level guard returns HHS_ERR_DEPTH (from
    the host recursion or from a device subtree)."""
    pos, vel, m = plummer(hhs, 16, seed=3)
    pos[5] = pos[4]
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(16, 0.01, 0.0, max_level=12, subtree_max=subtree_max)
    with pytest.raises(hhs.HHSError) as ei:
        g.hold_step(d, 1.0 / 64)
    assert ei.value.code == hhs.HHS_ERR_DEPTH


@pytest.mark.parametrize("n,eps", [(2, 1e-3), (5, 0.0), (64, 1e-3), (65, 0.0), (200, 1e-3),
                                   (256, 1e-4), (300, 1e-3), (700, 0.0), (1100, 1e-4),
                                   (2500, 1e-3)])
def test_subtree_fp32_timesteps_match_oracle(hhs, oracle, torch_cuda, n, eps):
    """FP32 timesteps of the device engine (each unordered pair once, column
    minima by atomics) against the FP64 oracle, within 1e-6 relative: a root
    step whose dt is below every timestep runs one call (all particles
    slow), so the timesteps left behind are those of the root call.  Up to
    256 particles the particles sit in LDS and each lane holds one sink;
    above, global memory and several sinks per lane and passes."""
    eta = 0.01
    pos, vel, m = plummer(hhs, n, seed=6060 + n, eps2=eps * eps)
    ref = oracle.timesteps(pos, vel, m, eta, eps * eps)
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(n, eta, eps, max_level=40, subtree_max=4096,
                    fast_timesteps=hhs.TIMESTEP_FP32)
    g.hold_step(d, float(ref.min()) * 0.5)
    torch_cuda.cuda.synchronize()
    assert g.stats()["calls"] == 1
    ts = d.timestep.cpu().numpy()
    assert np.all(np.abs(ts - ref) <= 1e-6 * ref)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("n", [37, 100, 256])
def test_hold_lds_subtree_bits_equal_ensemble(hhs, torch_cuda, n, mode):
    """A root step run as one device subtree (n <= 256: the particles live in
    LDS, a 256-thread team that spreads a call's sources over its four waves
    when it has <= 128 sinks) gives the same bits as the ensemble engine on
    the same system (particles in global memory, a one-wave team that forms
    the same source quarters one after the other), in every timestep mode."""
    eps, eta, dt = 1e-3, 0.01, 1.0 / 16
    pos, vel, m = plummer(hhs, n, seed=5150 + n, eps2=eps * eps)
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    # max_level 40 leaves the root call itself inside the subtree's depth range
    g = hhs.Globals(n, eta, eps, max_level=40, subtree_max=256, fast_timesteps=mode)
    g.hold(d, dt, dt)
    ens = hhs.Ensemble([(pos, vel, m)], dt, dt, eta, eps, max_level=40, fast_timesteps=mode)
    ens.run()
    torch_cuda.cuda.synchronize()
    st = g.stats()
    assert st["calls"] == ens.results()[0]["calls"] and st["calls"] > 1
    p, v = ens.state(0)
    assert d.pos.cpu().numpy().tobytes() == p.tobytes()
    assert d.vel.cpu().numpy().tobytes() == v.tobytes()


@pytest.mark.parametrize("subtree_max", [32, 256])
def test_hold_subtrees_several_root_steps(hhs, oracle, torch_cuda, subtree_max):
    """Device subtrees inside a longer run: same recursion and simtime as
    the FP64 CPU HOLD, trajectories within 1e-6 of the rms, and
    momentum conserved."""
    n, eps, eta = 400, 1e-3, 0.01
    pos, vel, m = plummer(hhs, n, seed=4646, eps2=eps * eps)
    host, ost, d, g = _run_pair(hhs, oracle, pos, vel, m, eta, eps, 0.125, 1.0 / 32,
                                subtree_max)
    torch_cuda.cuda.synchronize()
    st = g.stats()
    assert g.subtree_max == subtree_max
    assert st["calls"] == ost.calls
    assert st["calls_per_level"] == [int(x) for x in ost.calls_per_level[:ost.max_level + 1]]
    assert g.simtime == ost.simtime
    p, v = d.pos.cpu().numpy(), d.vel.cpu().numpy()
    assert np.abs(p - host.pos).max() <= 1e-6 * np.sqrt((host.pos**2).mean())
    assert np.abs(v - host.vel).max() <= 1e-6 * np.sqrt((host.vel**2).mean())
    mom = (m[:, None] * v).sum(0)
    assert np.abs(mom).max() <= 1e-6 * (m[:, None] * np.abs(v)).sum(0).max()


@pytest.mark.parametrize("n", [64, 300])
def test_hold_fp32_sets_track_oracle(hhs, oracle, torch_cuda, n):
    """hold_step on FP32 particle sets (the reference's T = float path,
    src/hold.cuh:51-53): the same recursion as the FP64 CPU HOLD started from
    the FP32-rounded state, positions and velocities within 1e-5 of the rms
    (FP32 storage), momentum conserved to FP32 rounding."""
    import torch
    eps, eta, dt = 1e-2, 0.01, 1.0 / 32
    pos, vel, m = plummer(hhs, n, seed=9100 + n, eps2=eps * eps)
    pos, vel, m = (a.astype(np.float32).astype(np.float64) for a in (pos, vel, m))
    host = oracle.HostParticles(n, pos, vel, m)
    rc, ost = oracle.hold(host, dt, dt, eta, eps * eps)
    assert rc == 0
    d = hhs.DeviceParticles.from_host(pos, vel, m, dtype=torch.float32)
    g = hhs.Globals(n, eta, eps)
    g.hold(d, dt, dt)
    torch_cuda.cuda.synchronize()
    st = g.stats()
    assert st["calls"] == ost.calls and st["interactions"] == ost.interactions
    assert g.simtime == ost.simtime
    p = d.pos.cpu().numpy().astype(np.float64)
    v = d.vel.cpu().numpy().astype(np.float64)
    assert np.abs(p - host.pos).max() <= 1e-5 * np.sqrt((host.pos**2).mean())
    assert np.abs(v - host.vel).max() <= 1e-5 * np.sqrt((host.vel**2).mean())
    mom = (m[:, None] * v).sum(0)
    assert np.abs(mom).max() <= 1e-5 * (m[:, None] * np.abs(v)).sum(0).max()


@pytest.mark.parametrize("mode", ["fp64", "default_host", "default_subtree"])
def test_hold_calc_timestep_flags_closed_form(hhs, torch_cuda, mode):
    """The hand-built triple whose call tree follows in closed form (the
    first recursion reuses the timesteps, the second recomputes them,
    src/hold.cuh:116-121,168-173): calls per level [1, 2, 2] and 10
    timestep pairs on the host recursion (FP64-only and default kicks) and
    in a device subtree."""
    from conftest import flags_triple
    pos, vel, m, eta, dt, _ = flags_triple()
    kw = {"fp64": dict(kick_precision=hhs.KICK_F64_EXACT),
          "default_host": dict(fast_timesteps=hhs.TIMESTEP_FP32, subtree_max=0),
          "default_subtree": dict(fast_timesteps=hhs.TIMESTEP_FP32, max_level=40,
                                  subtree_max=256)}[mode]
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(3, eta, 0.0, **kw)
    g.hold_step(d, dt)
    torch_cuda.cuda.synchronize()
    st = g.stats()
    assert st["calls_per_level"] == [1, 2, 2] and st["ts_pairs"] == 10 and st["calls"] == 5
    assert g.simtime == dt


@pytest.mark.parametrize("case", ["default", "no_subtree", "subtree512", "fp32_sets",
                                  "fp64_exact", "tiny"])
def test_hold_speculation_same_bits(hhs, torch_cuda, case):
    """Speculative launches (each call's first fast recursion queued before
    its counts are read back, with the set size read on the device) change
    nothing: positions, velocities, statistics and simtime are the same bits
    with speculation on and off, over several root steps, with and without
    device subtrees, with subtree_max above the speculative subtree's limit,
    on FP32 sets and in the FP64-only mode."""
    import torch
    n, kw, dtype = {
        "default": (3000, dict(fast_timesteps=hhs.TIMESTEP_FP32), torch.float64),
        "no_subtree": (1500, dict(fast_timesteps=hhs.TIMESTEP_FP32, subtree_max=0),
                       torch.float64),
        "subtree512": (1500, dict(subtree_max=512), torch.float64),
        "fp32_sets": (1200, dict(), torch.float32),
        "fp64_exact": (600, dict(kick_precision=hhs.KICK_F64_EXACT), torch.float64),
        "tiny": (16, dict(subtree_max=0), torch.float64),
    }[case]
    eps, eta = 1e-3, 0.01
    pos, vel, m = plummer(hhs, n, seed=7300 + n, eps2=eps * eps)
    out = []
    for spec in (True, False):
        d = hhs.DeviceParticles.from_host(pos, vel, m, dtype=dtype)
        g = hhs.Globals(n, eta, eps, max_level=40, speculation=spec, **kw)
        g.hold(d, 0.125, 1.0 / 16)
        torch_cuda.cuda.synchronize()
        out.append((d.pos.cpu().numpy().tobytes(), d.vel.cpu().numpy().tobytes(), g.stats(),
                    g.simtime))
    assert out[0][2]["calls"] > 3
    assert out[0] == out[1]


@pytest.mark.parametrize("case", ["default", "no_speculation", "eps0", "root_step_16k"])
def test_hold_fused_reduce_same_bits(hhs, torch_cuda, case):
    """A call's split-K kick partials added by their readers -- the device
    subtree of its second fast recursion (fast sinks) and its scatter-back
    (slow sinks) -- instead of a reduce launch: positions, velocities,
    timesteps, statistics and simtime are the same bits as with the reduce
    launch, over several root steps; also with eps = 0 (unsoftened kicks),
    without speculative launches, and on one root step of the N = 16384
    sphere."""
    n, eps, t_end, dt = {"default": (3000, 1e-3, 0.125, 1.0 / 16),
                         "no_speculation": (2000, 1e-3, 0.125, 1.0 / 16),
                         "eps0": (1500, 0.0, 0.125, 1.0 / 16),
                         "root_step_16k": (16384, 1e-4, 0.125, 0.125)}[case]
    kw = dict(fast_timesteps=hhs.TIMESTEP_FP32)
    if case == "no_speculation":
        kw["speculation"] = False
    pos, vel, m = plummer(hhs, n, seed=7500 + n, eps2=eps * eps)
    out = []
    for fused in (True, False):
        d = hhs.DeviceParticles.from_host(pos, vel, m)
        g = hhs.Globals(n, 0.01, eps, max_level=40, fused_reduce=fused, **kw)
        g.hold(d, t_end, dt)
        torch_cuda.cuda.synchronize()
        out.append((d.pos.cpu().numpy().tobytes(), d.vel.cpu().numpy().tobytes(),
                    d.timestep.cpu().numpy().tobytes(), g.stats(), g.simtime))
        g.close()
    assert out[0][3]["calls"] > 3
    assert out[0] == out[1]


@pytest.mark.parametrize("case", ["default", "no_subtree", "fp32_sets", "no_fused_reduce",
                                  "fp64_exact", "root_step_16k"])
def test_hold_chain_scatter_same_bits(hhs, torch_cuda, case):
    """The scatter-backs of a chain of calls, each its parent's second fast
    recursion, run as one launch: positions, velocities, timesteps,
    statistics and simtime are the same bits as with one scatter-back launch
    per call, over several root steps; with and without device subtrees
    (the host recursion then reaches the deepest levels, and chains longer
    than one launch holds), on FP32 sets, without the fused kick reduce, in
    the FP64-only mode, and on one root step of the N = 16384 sphere."""
    import torch
    n, eps, t_end, dt = 2500, 1e-3, 0.125, 1.0 / 16
    kw, dtype = dict(fast_timesteps=hhs.TIMESTEP_FP32), torch.float64
    if case == "no_subtree":
        kw["subtree_max"] = 0
    elif case == "fp32_sets":
        kw, dtype = {}, torch.float32
    elif case == "no_fused_reduce":
        kw["fused_reduce"] = False
    elif case == "fp64_exact":
        n, kw = 600, dict(kick_precision=hhs.KICK_F64_EXACT)
    elif case == "root_step_16k":
        n, eps, t_end, dt = 16384, 1e-4, 0.125, 0.125
    pos, vel, m = plummer(hhs, n, seed=7600 + n, eps2=eps * eps)
    out = []
    for chain in (True, False):
        d = hhs.DeviceParticles.from_host(pos, vel, m, dtype=dtype)
        g = hhs.Globals(n, 0.01, eps, max_level=40, chain_scatter=chain, **kw)
        g.hold(d, t_end, dt)
        torch_cuda.cuda.synchronize()
        out.append((d.pos.cpu().numpy().tobytes(), d.vel.cpu().numpy().tobytes(),
                    d.timestep.cpu().numpy().tobytes(), g.stats(), g.simtime))
        g.close()
    assert out[0][3]["calls"] > 3
    assert out[0] == out[1]
