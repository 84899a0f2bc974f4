"""HOLD on the benchmark configurations, and the decision guard of the
approximate timestep modes.

The FP32 (and FAST) timestep modes recompute, with the exact formula, every
timestep that lies within the guard band (1e-5 relative in FP32) of a pivot
dt * 2^-k, so their slow/fast decisions are the exact mode's.  Timesteps
enter HOLD only through those decisions (src/split.cuh:115,
src/hold.cuh:79-83), so with the same kick arithmetic an FP32-timestep run
and an exact-timestep run must give the same bits -- the tests below check
exactly that, on the host recursion (exact kicks) and on the device engine
(subtrees and ensembles), including inputs built so that the FP32 and exact
timesteps of one particle straddle the pivot.

Configurations (the bars of the north star):
* N = 16384 equal-mass Plummer sphere (seed 2, eps = 1e-4, eta = 0.01) to
  t = 0.125 in one root step;
* N = 131072 Plummer sphere with Salpeter masses (seed 3, eps = 1e-4,
  eta = 0.01), one root step of 1/64;
the default mode (MIXED32 kicks, FP32 timesteps, device subtrees) against
the FP64-only mode (F64_EXACT kicks, exact timesteps; bit-identical with the
FP64 CPU HOLD): identical calls per level, interactions and timestep pairs,
|dE/E| within 10 % of the FP64 run's, positions and velocities within 1e-6
of the rms.
"""
import numpy as np
import pytest

from conftest import plummer

pytestmark = pytest.mark.gpu

GUARD_BAND_FP32 = 1e-5


def _ts(hhs, torch_cuda, pos, vel, m, eta, eps2, mode, sym=False):
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    if sym:
        hhs.timesteps_sym(d, eta=eta, eps2=eps2, mode=mode)
    else:
        hhs.timesteps(d, eta=eta, eps2=eps2, fast=mode)
    torch_cuda.cuda.synchronize()
    return d.timestep.cpu().numpy()


@pytest.mark.parametrize("n,seed,mass_model,sym", [(16384, 2, 0, False), (131072, 3, 1, False),
                                                   (131072, 3, 1, True)])
def test_fp32_timestep_error_inside_guard_band(hhs, torch_cuda, n, seed, mass_model, sym):
    """The guard's premise: on the benchmark ICs the FP32 timesteps (one-sided
    and symmetric kernels) stay well inside the band, at most a tenth of it."""
    eps, eta = 1e-4, 0.01
    pos, vel, m = plummer(hhs, n, seed=seed, mass_model=mass_model, eps2=eps * eps)
    ref = _ts(hhs, torch_cuda, pos, vel, m, eta, eps * eps, hhs.TIMESTEP_EXACT, sym)
    t32 = _ts(hhs, torch_cuda, pos, vel, m, eta, eps * eps, hhs.TIMESTEP_FP32, sym)
    rel = np.abs(t32 / ref - 1).max()
    print(f"n={n} sym={sym}: max |tau32/tau64 - 1| = {rel:.3e}")
    assert rel <= GUARD_BAND_FP32 / 10


def _pair_runs(hhs, pos, vel, m, eta, eps, dt, t_end, modes, kick, subtree_max=None):
    out = []
    for mode in modes:
        d = hhs.DeviceParticles.from_host(pos, vel, m)
        g = hhs.Globals(len(m), eta, eps, fast_timesteps=mode, kick_precision=kick,
                        subtree_max=subtree_max)
        g.hold(d, t_end, dt)
        out.append((d, g))
    return out


def _assert_same_bits(runs):
    (d0, g0), (d1, g1) = runs
    assert g0.stats() == g1.stats()
    assert g0.simtime == g1.simtime
    assert d0.pos.cpu().numpy().tobytes() == d1.pos.cpu().numpy().tobytes()
    assert d0.vel.cpu().numpy().tobytes() == d1.vel.cpu().numpy().tobytes()


@pytest.mark.parametrize("mode", [2, 1])
@pytest.mark.parametrize("n,eps", [(1024, 0.0), (3000, 1e-4)])
def test_guard_host_recursion_equals_exact(hhs, torch_cuda, n, eps, mode):
    """Host recursion with exact kicks: FP32 (mode 2) or FAST (mode 1)
    timesteps with the guard give the bits of exact timesteps over several
    root steps."""
    eta = 0.01
    pos, vel, m = plummer(hhs, n, seed=7100 + n, eps2=eps * eps)
    runs = _pair_runs(hhs, pos, vel, m, eta, eps, 1.0 / 32, 1.0 / 8, [mode, hhs.TIMESTEP_EXACT],
                      hhs.KICK_F64_EXACT)
    torch_cuda.cuda.synchronize()
    _assert_same_bits(runs)


def test_guard_stage_after_symmetric_equals_exact(hhs, torch_cuda):
    """From 8192 particles the FP32 timesteps come from the symmetric kernel,
    and its guard stage (flagged list, per-workgroup exact partials, the last
    workgroup's minima) replaces the near-pivot timesteps: with exact kicks
    the run gives the bits of exact timesteps over several root steps."""
    n, eps, eta = 9000, 1e-4, 0.01
    pos, vel, m = plummer(hhs, n, seed=7150, eps2=eps * eps)
    runs = _pair_runs(hhs, pos, vel, m, eta, eps, 1.0 / 64, 1.0 / 32,
                      [hhs.TIMESTEP_FP32, hhs.TIMESTEP_EXACT], hhs.KICK_F64_EXACT)
    torch_cuda.cuda.synchronize()
    _assert_same_bits(runs)
    # a root dt between one particle's FP32 and exact timesteps, so that the
    # stage has at least that particle to fix in the root call
    t64 = _ts(hhs, torch_cuda, pos, vel, m, eta, eps * eps, hhs.TIMESTEP_EXACT)
    t32 = _ts(hhs, torch_cuda, pos, vel, m, eta, eps * eps, hhs.TIMESTEP_FP32, True)
    dt, i = _straddle_dt(t32, t64)
    assert (t32[i] <= dt) != (t64[i] <= dt)
    runs = _pair_runs(hhs, pos, vel, m, eta, eps, dt, dt, [hhs.TIMESTEP_FP32, hhs.TIMESTEP_EXACT],
                      hhs.KICK_F64_EXACT)
    torch_cuda.cuda.synchronize()
    _assert_same_bits(runs)


def _straddle_dt(t32, t64):
    """A dt between the FP32 and exact timesteps of the particle where they
    differ most (relative), so that an unguarded FP32 split decides that
    particle differently from the exact one."""
    rel = np.abs(t32 / t64 - 1)
    i = int(np.argmax(rel))
    assert rel[i] > 0, "FP32 and exact timesteps agree everywhere"
    return 0.5 * (t32[i] + t64[i]), i


@pytest.mark.parametrize("n", [600, 4000])
def test_guard_host_split_at_straddling_pivot(hhs, torch_cuda, n):
    """One root step whose dt lies between a particle's FP32 and exact
    timesteps: the guarded FP32 run decides it as the exact run does (same
    recursion and bits; exact kicks, host recursion)."""
    eps, eta = 1e-3, 0.01
    pos, vel, m = plummer(hhs, n, seed=7300 + n, eps2=eps * eps)
    t64 = _ts(hhs, torch_cuda, pos, vel, m, eta, eps * eps, hhs.TIMESTEP_EXACT)
    t32 = _ts(hhs, torch_cuda, pos, vel, m, eta, eps * eps, hhs.TIMESTEP_FP32)
    dt, i = _straddle_dt(t32, t64)
    assert (t32[i] <= dt) != (t64[i] <= dt)
    runs = _pair_runs(hhs, pos, vel, m, eta, eps, dt, dt, [hhs.TIMESTEP_FP32, hhs.TIMESTEP_EXACT],
                      hhs.KICK_F64_EXACT)
    torch_cuda.cuda.synchronize()
    _assert_same_bits(runs)


@pytest.mark.parametrize("n", [200, 1024])
def test_guard_subtrees_equal_exact(hhs, torch_cuda, n):
    """Default kicks with device subtrees (n = 200: the whole root step runs
    in one subtree; n = 1024: host calls above 256 particles, subtrees
    below): FP32 timesteps with the guard give the bits of exact timesteps."""
    eps, eta = 1e-4, 0.01
    pos, vel, m = plummer(hhs, n, seed=7500 + n, eps2=eps * eps)
    runs = _pair_runs(hhs, pos, vel, m, eta, eps, 1.0 / 16, 1.0 / 8,
                      [hhs.TIMESTEP_FP32, hhs.TIMESTEP_EXACT], hhs.KICK_MIXED32, subtree_max=256)
    torch_cuda.cuda.synchronize()
    _assert_same_bits(runs)


def test_guard_subtree_at_straddling_pivot(hhs, torch_cuda):
    """The engine's own FP32 timesteps (a root step below every timestep
    leaves them behind) against exact ones, and a root step whose dt lies
    between them for one particle: FP32 and exact runs give the same bits."""
    n, eps, eta = 180, 1e-3, 0.01
    pos, vel, m = plummer(hhs, n, seed=7700, eps2=eps * eps)
    t64 = _ts(hhs, torch_cuda, pos, vel, m, eta, eps * eps, hhs.TIMESTEP_EXACT)
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    g = hhs.Globals(n, eta, eps, max_level=40, subtree_max=256, fast_timesteps=hhs.TIMESTEP_FP32)
    g.hold_step(d, float(t64.min()) * 0.25)
    torch_cuda.cuda.synchronize()
    assert g.stats()["calls"] == 1
    t32 = d.timestep.cpu().numpy()
    dt, i = _straddle_dt(t32, t64)
    assert (t32[i] <= dt) != (t64[i] <= dt)
    out = []
    for mode in (hhs.TIMESTEP_FP32, hhs.TIMESTEP_EXACT):
        d = hhs.DeviceParticles.from_host(pos, vel, m)
        g = hhs.Globals(n, eta, eps, max_level=40, subtree_max=256, fast_timesteps=mode)
        g.hold_step(d, dt)
        out.append((d, g))
    torch_cuda.cuda.synchronize()
    _assert_same_bits(out)


def test_guard_ensemble_equals_exact(hhs, torch_cuda):
    """A batch of systems (wave teams and workgroup teams) with FP32
    timesteps gives the bits of the same batch with exact timesteps."""
    eps, eta = 1e-4, 0.01
    sizes = [8, 30, 64, 150, 300, 513, 900, 1500]
    systems = [plummer(hhs, n, seed=7900 + k, eps2=eps * eps) for k, n in enumerate(sizes)]
    res = []
    for mode in (hhs.TIMESTEP_FP32, hhs.TIMESTEP_EXACT):
        ens = hhs.Ensemble(systems, 0.25, 1.0 / 16, eta, eps, fast_timesteps=mode)
        ens.run()
        torch_cuda.cuda.synchronize()
        res.append(ens)
    a, b = res
    for k in range(len(sizes)):
        ra, rb = a.results()[k], b.results()[k]
        assert ra["status"] == 0 and rb["status"] == 0
        for key in ("calls", "interactions", "max_level", "e1"):
            assert ra[key] == rb[key], (sizes[k], key)
        pa, va = a.state(k)
        pb, vb = b.state(k)
        assert pa.tobytes() == pb.tobytes() and va.tobytes() == vb.tobytes()


def _energy(hhs, d, eps):
    ke, pe = hhs.energy(d.pos, d.vel, d.m, eps * eps)
    return ke + pe


def _config_vs_fp64(hhs, torch_cuda, n, seed, mass_model, t_end):
    eps, eta = 1e-4, 0.01
    pos, vel, m = plummer(hhs, n, seed=seed, mass_model=mass_model, eps2=eps * eps)
    runs = {}
    for name, ts_mode, kick in (("default", hhs.TIMESTEP_FP32, hhs.KICK_MIXED32),
                                ("fp64", hhs.TIMESTEP_EXACT, hhs.KICK_F64_EXACT)):
        d = hhs.DeviceParticles.from_host(pos, vel, m)
        e0 = _energy(hhs, d, eps)
        g = hhs.Globals(n, eta, eps, fast_timesteps=ts_mode, kick_precision=kick)
        g.hold_step(d, t_end)
        torch_cuda.cuda.synchronize()
        de = abs(_energy(hhs, d, eps) / e0 - 1)
        runs[name] = (d, g, de)
    (d0, g0, de0), (d1, g1, de1) = runs["default"], runs["fp64"]
    s0, s1 = g0.stats(), g1.stats()
    print(f"n={n}: calls {s0['calls']}/{s1['calls']}, |dE/E| {de0:.4e} vs fp64 {de1:.4e} "
          f"(ratio {de0 / de1:.4f})")
    assert s0 == s1
    assert g0.simtime == g1.simtime
    assert abs(de0 - de1) <= 0.10 * de1
    for f in ("pos", "vel"):
        a, b = getattr(d0, f).cpu().numpy(), getattr(d1, f).cpu().numpy()
        assert np.abs(a - b).max() <= 1e-6 * np.sqrt((b**2).mean())


def test_hold_n16384_default_matches_fp64_run(hhs, torch_cuda):
    _config_vs_fp64(hhs, torch_cuda, 16384, 2, 0, 0.125)


def test_hold_n131072_root_step_default_matches_fp64_run(hhs, torch_cuda):
    _config_vs_fp64(hhs, torch_cuda, 131072, 3, 1, 1.0 / 64)


def test_guard_stage_overflow_on_a_lattice(hhs, torch_cuda):
    """The symmetric kernel's guard stage takes at most 256 flagged
    timesteps; beyond that the split runs its own guard.  A cubic lattice at
    rest gives every particle the same free-fall timestep; with the root
    step just above it, all 16384 timesteps lie in the band (the stage
    overflows), and the FP32 run must still make the exact run's decisions:
    the same bits."""
    nx, ny, nz, a = 16, 32, 32, 1.0 / 32
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    pos = (g.reshape(-1, 3) * a).astype(np.float64)
    n = len(pos)
    vel = np.zeros_like(pos)
    m = np.full(n, 1.0 / n)
    eps, eta = 1e-4, 0.01
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    hhs.timesteps(d, eta=eta, eps2=eps * eps, fast=hhs.TIMESTEP_EXACT)
    torch_cuda.cuda.synchronize()
    t64 = d.timestep.cpu().numpy()
    assert np.all(t64 == t64[0])
    dt = float(t64[0]) * (1 + 2e-6)
    out = []
    for mode in (hhs.TIMESTEP_FP32, hhs.TIMESTEP_EXACT):
        dd = hhs.DeviceParticles.from_host(pos, vel, m)
        gg = hhs.Globals(n, eta, eps, fast_timesteps=mode)
        gg.hold_step(dd, dt)
        out.append((dd, gg))
    torch_cuda.cuda.synchronize()
    (d0, g0), (d1, g1) = out
    print("lattice stats", g0.stats()["calls"], g1.stats()["calls"])
    _assert_same_bits(out)
