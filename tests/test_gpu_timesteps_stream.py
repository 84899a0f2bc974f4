"""Timesteps, drift, update and energy: parity with the FP64 oracle.

Timesteps, drift and update use the oracle's operation order with IEEE
sqrt/divide and no FMA contraction, so they are compared bit for bit.
Energy is compared to 1e-12 relative (the final sum order differs).
"""
import numpy as np
import pytest

from conftest import plummer

pytestmark = pytest.mark.gpu


# Below ~32768 particles the sources are split over several workgroups and
# the chunk minima merged in order (split-K); the minimum stays exact.
@pytest.mark.parametrize("n", [2, 3, 64, 65, 257, 1024, 3000, 5001])
@pytest.mark.parametrize("eps2", [0.0, 1e-8])
def test_timesteps_bit_exact(hhs, oracle, torch_cuda, n, eps2):
    pos, vel, m = plummer(hhs, n, seed=40 + n, mass_model=1, eps2=eps2)
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    hhs.timesteps(d, eta=0.01, eps2=eps2)
    torch_cuda.cuda.synchronize()
    ref = oracle.timesteps(pos, vel, m, 0.01, eps2)
    got = d.timestep.cpu().numpy()
    assert np.array_equal(got, ref), f"max rel diff {np.max(np.abs(got / ref - 1)):.3e}"


def test_timesteps_single_particle_and_cold_start(hhs, oracle, torch_cuda):
    d = hhs.DeviceParticles.from_host(np.zeros((1, 3)), np.zeros((1, 3)), np.ones(1))
    hhs.timesteps(d, eta=0.01)
    torch_cuda.cuda.synchronize()
    assert d.timestep.cpu().numpy()[0] == 1e30  # src/timestep.cuh:4 start value
    # cold start: all velocities zero -> fly-by candidates are NaN/inf and never win
    pos, _, m = plummer(hhs, 100, seed=5)
    vel = np.zeros((100, 3))
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    hhs.timesteps(d, eta=0.02)
    torch_cuda.cuda.synchronize()
    assert np.array_equal(d.timestep.cpu().numpy(), oracle.timesteps(pos, vel, m, 0.02))


def test_timesteps_circular_pair_closed_form(hhs, torch_cuda):
    # v . r = 0: tau_ff = eta/sqrt2 * sqrt(r^3/M), tau_fb = eta * r / |dv|
    r, M, eta = 0.5, 1.0, 0.03
    v = np.sqrt(M / r)
    pos = np.array([[-r / 2, 0, 0], [r / 2, 0, 0]])
    vel = np.array([[0, -v / 2, 0], [0, v / 2, 0]])
    m = np.array([0.5, 0.5])
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    hhs.timesteps(d, eta=eta)
    torch_cuda.cuda.synchronize()
    tff = eta / np.sqrt(2) * np.sqrt(r**3 / M)
    tfb = eta * r / v
    got = d.timestep.cpu().numpy()
    assert np.allclose(got, min(tff, tfb), rtol=1e-14)


@pytest.mark.parametrize("n", [2, 65, 1024, 3000])
@pytest.mark.parametrize("eps2", [0.0, 1e-8])
def test_timesteps_fast_within_few_ulp(hhs, oracle, torch_cuda, n, eps2):
    """FAST mode (rsqrt + Newton, fractions compared by cross-multiplication)
    against the IEEE oracle: relative difference <= 1e-14 (a few ulp)."""
    pos, vel, m = plummer(hhs, n, seed=90 + n, mass_model=1, eps2=eps2)
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    hhs.timesteps(d, eta=0.01, eps2=eps2, fast=True)
    torch_cuda.cuda.synchronize()
    ref = oracle.timesteps(pos, vel, m, 0.01, eps2)
    got = d.timestep.cpu().numpy()
    assert np.max(np.abs(got / ref - 1.0)) <= 1e-14


@pytest.mark.parametrize("n", [2, 65, 1024, 3000, 20000])
@pytest.mark.parametrize("eps2", [0.0, 1e-8])
def test_timesteps_fp32_within_1e6(hhs, oracle, torch_cuda, n, eps2):
    """FP32 mode (FP64 differences rounded once to FP32): relative
    difference <= 1e-6 against the FP64 oracle (the timestep bar)."""
    pos, vel, m = plummer(hhs, n, seed=190 + n, mass_model=1, eps2=eps2)
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    hhs.timesteps(d, eta=0.01, eps2=eps2, fast=hhs.TIMESTEP_FP32)
    torch_cuda.cuda.synchronize()
    if n <= 3000:
        ref = oracle.timesteps(pos, vel, m, 0.01, eps2)
        got = d.timestep.cpu().numpy()
    else:  # FP64 GPU exact mode is bit-exact with the oracle (tested above)
        e = hhs.DeviceParticles.from_host(pos, vel, m)
        hhs.timesteps(e, eta=0.01, eps2=eps2)
        torch_cuda.cuda.synchronize()
        ref, got = e.timestep.cpu().numpy(), d.timestep.cpu().numpy()
    assert np.max(np.abs(got / ref - 1.0)) <= 1e-6


def test_timesteps_fast_edge_cases(hhs, oracle, torch_cuda):
    # single particle: start value; cold start: fly-by terms never win;
    # coincident pair with eps = 0: tau = 0 exactly (as the IEEE form)
    d = hhs.DeviceParticles.from_host(np.zeros((1, 3)), np.zeros((1, 3)), np.ones(1))
    hhs.timesteps(d, eta=0.01, fast=True)
    torch_cuda.cuda.synchronize()
    assert d.timestep.cpu().numpy()[0] == 1e30
    pos, _, m = plummer(hhs, 100, seed=5)
    vel = np.zeros((100, 3))
    pos[7] = pos[3]
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    hhs.timesteps(d, eta=0.02, fast=True)
    torch_cuda.cuda.synchronize()
    ref = oracle.timesteps(pos, vel, m, 0.02)
    got = d.timestep.cpu().numpy()
    assert got[3] == 0.0 and got[7] == 0.0 and ref[3] == 0.0
    ok = ref > 0
    assert np.max(np.abs(got[ok] / ref[ok] - 1.0)) <= 1e-14


@pytest.mark.parametrize("n", [1, 100003, 100004])
def test_drift_bit_exact(hhs, oracle, torch_cuda, n):
    """Vector path (16-byte loads) with an odd tail of components (3n odd),
    and an even count."""
    torch = torch_cuda
    rng = np.random.default_rng(0)
    pos, vel = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    p = torch.as_tensor(pos).cuda()
    v = torch.as_tensor(vel).cuda()
    hhs.drift(p, v, dt=0.0123, fac=0.5)
    torch.cuda.synchronize()
    assert np.array_equal(p.cpu().numpy(), oracle.drift(pos, vel, 0.0123, 0.5))


def test_drift_unaligned_and_fp32(hhs, oracle, torch_cuda):
    """Arrays that start 8 bytes past a 16-byte boundary take the scalar
    path; FP32 arrays the four-component vector path with a tail.  Both give
    the element expression pos + vel * h (FP64 bit-exact; FP32 against the
    same expression in FP32)."""
    torch = torch_cuda
    n = 5001
    rng = np.random.default_rng(3)
    pos, vel = rng.standard_normal((n + 1, 3)), rng.standard_normal((n + 1, 3))
    p = torch.as_tensor(pos).cuda().reshape(-1)
    v = torch.as_tensor(vel).cuda().reshape(-1)
    ps, vs = p[1:1 + 3 * n], v[1:1 + 3 * n]   # 8 bytes past the allocation's start
    hhs.drift(ps, vs, n=n, dt=0.0123, fac=0.5)
    torch.cuda.synchronize()
    ref = pos.reshape(-1).copy()
    ref[1:1 + 3 * n] = oracle.drift(pos.reshape(-1)[1:1 + 3 * n].reshape(n, 3),
                                    vel.reshape(-1)[1:1 + 3 * n].reshape(n, 3), 0.0123,
                                    0.5).reshape(-1)
    assert np.array_equal(p.cpu().numpy(), ref)
    p32, v32 = pos[:n].astype(np.float32), vel[:n].astype(np.float32)
    d32 = torch.as_tensor(p32).cuda()
    hhs.drift(d32, torch.as_tensor(v32).cuda(), dt=0.0123, fac=0.5)
    torch.cuda.synchronize()
    h = np.float32(0.0123) * np.float32(0.5)
    assert np.array_equal(d32.cpu().numpy(), p32 + v32 * h)


def test_update_inverse_permutation(hhs, oracle, torch_cuda):
    torch = torch_cuda
    n = 70001
    rng = np.random.default_rng(1)
    perm = rng.permutation(n).astype(np.int32)
    cur = hhs.DeviceParticles.from_host(rng.standard_normal((n, 3)), rng.standard_normal((n, 3)),
                                        rng.random(n), rng.random(n))
    cur.parent_id.copy_(torch.as_tensor(perm))
    prev = hhs.DeviceParticles.empty(n)
    hhs.update(prev, cur)
    torch.cuda.synchronize()
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    for f in ("pos", "vel", "m", "timestep"):
        assert np.array_equal(getattr(prev, f).cpu().numpy(), getattr(cur, f).cpu().numpy()[inv])


@pytest.mark.parametrize("n,eps2", [(2, 0.0), (1000, 0.0), (5000, 1e-8)])
def test_energy_matches_oracle(hhs, oracle, torch_cuda, n, eps2):
    pos, vel, m = plummer(hhs, n, seed=n, eps2=eps2)
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    ke, pe = hhs.energy(d.pos, d.vel, d.m, eps2)
    assert abs(ke - oracle.kinetic(vel, m)) <= 1e-12 * abs(ke)
    assert abs(pe - oracle.potential(pos, m, eps2)) <= 1e-12 * abs(pe)


@pytest.mark.parametrize("n", [1, 2, 3, 65, 1024, 1025, 3000])
@pytest.mark.parametrize("eps2", [0.0, 1e-8])
def test_timesteps_symmetric_bit_exact(hhs, oracle, torch_cuda, n, eps2):
    """Symmetric evaluation (each unordered pair once, partial minima then a
    min-reduction): EXACT mode bit-exact with the oracle, including ragged
    last blocks (1025) and a coincident pair with eps2 = 0."""
    pos, vel, m = plummer(hhs, n, seed=640 + n, mass_model=1, eps2=eps2)
    if n > 10:
        pos[9] = pos[4]
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    hhs.timesteps_sym(d, eta=0.01, eps2=eps2, mode=hhs.TIMESTEP_EXACT)
    torch_cuda.cuda.synchronize()
    ref = oracle.timesteps(pos, vel, m, 0.01, eps2)
    assert d.timestep.cpu().numpy().tobytes() == ref.tobytes()


@pytest.mark.parametrize("n", [3000, 40000])
def test_timesteps_symmetric_large(hhs, oracle, torch_cuda, n):
    """At sizes HOLD runs it: EXACT equals the one-sided EXACT kernel (itself
    bit-exact with the oracle) bit for bit; FP32 within 1e-6 of it."""
    eps2 = 1e-8
    pos, vel, m = plummer(hhs, n, seed=700 + n, mass_model=1, eps2=eps2)
    a = hhs.DeviceParticles.from_host(pos, vel, m)
    b = hhs.DeviceParticles.from_host(pos, vel, m)
    c = hhs.DeviceParticles.from_host(pos, vel, m)
    hhs.timesteps(a, eta=0.01, eps2=eps2)
    hhs.timesteps_sym(b, eta=0.01, eps2=eps2, mode=hhs.TIMESTEP_EXACT)
    hhs.timesteps_sym(c, eta=0.01, eps2=eps2, mode=hhs.TIMESTEP_FP32)
    torch_cuda.cuda.synchronize()
    ta, tb, tc = (x.timestep.cpu().numpy() for x in (a, b, c))
    assert ta.tobytes() == tb.tobytes()
    assert np.max(np.abs(tc / ta - 1.0)) <= 1e-6


@pytest.mark.parametrize("n", [1, 63, 64, 65, 257, 1000, 4097, 9000, 16384, 20000])
def test_timesteps_sym_fp32_units_same_bits(hhs, torch_cuda, n):
    """The FP32 symmetric kernel's source unit (64, 128 or 256 sources per
    workgroup tile) and sinks per lane (3 or 4) change no bit of the
    timesteps: every unordered pair gives the same two values whichever
    particle is the sink, and a minimum does not depend on the grouping.
    Ragged sizes, a unit narrower than a wave, and the default choice."""
    eps2 = 1e-8
    pos, vel, m = plummer(hhs, n, seed=720 + n, mass_model=1, eps2=eps2)
    if n > 10:
        pos[9] = pos[4] + 1e-7   # a close pair
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    out = {}
    try:
        for unit in (0, 64, 128, 256):
            for spl in (3, 4):
                hhs.check(hhs._lib.hhs_timesteps_sym_set_unit(unit))
                hhs.check(hhs._lib.hhs_timesteps_sym_set_sinks_per_lane(spl))
                d.timestep.zero_()
                hhs.timesteps_sym(d, eta=0.01, eps2=eps2, mode=hhs.TIMESTEP_FP32)
                torch_cuda.cuda.synchronize()
                out[(unit, spl)] = d.timestep.cpu().numpy().tobytes()
    finally:
        hhs._lib.hhs_timesteps_sym_set_unit(0)
        hhs._lib.hhs_timesteps_sym_set_sinks_per_lane(0)
    ref = out[(256, 3)]
    assert all(v == ref for v in out.values())
    assert hhs._lib.hhs_timesteps_sym_set_unit(96) == hhs.HHS_ERR_INVALID


@pytest.mark.parametrize("mode", [0, 2])
def test_timestep_autotune_installs_a_valid_config(hhs, oracle, torch_cuda, tmp_path, mode):
    """The timestep autotuner installs a (W, K) for the size class and mode,
    writes it to the cache, and results keep their bar under it: EXACT bit
    for bit, FP32 within 1e-6."""
    path = tmp_path / "tune.txt"
    try:
        w, k, ms = hhs.autotune_timesteps(2048, mode, 2, str(path))
        assert w in (1, 2, 4, 8) and k in (1, 2, 4, 8, 16, 32, 64) and ms > 0
        assert hhs.timesteps_get_config(3000, mode) == (w, k)   # same size class (11)
        fields = path.read_text().split()
        assert fields[:5] == ["ts", str(mode), "11", str(w), str(k)]
        pos, vel, m = plummer(hhs, 3000, seed=77, eps2=1e-8)
        d = hhs.DeviceParticles.from_host(pos, vel, m)
        hhs.timesteps(d, eta=0.01, eps2=1e-8, fast=mode)
        torch_cuda.cuda.synchronize()
        ref = oracle.timesteps(pos, vel, m, 0.01, 1e-8)
        got = d.timestep.cpu().numpy()
        if mode == 0:
            assert got.tobytes() == ref.tobytes()
        else:
            assert np.all(np.abs(got - ref) <= 1e-6 * ref)
    finally:
        hhs.autotune_clear()


def test_energy_reports_nonfinite_state(hhs, torch_cuda):
    """A NaN in the state makes the FP64 energy diagnostic return
    HHS_ERR_NONFINITE instead of a NaN energy."""
    pos, vel, m = hhs.ic_plummer(300, 17, 0, True, 0.0)
    vel[123, 1] = np.nan
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    with pytest.raises(hhs.HHSError) as ei:
        hhs.energy(d.pos, d.vel, d.m, 1e-8)
    assert ei.value.code == hhs.HHS_ERR_NONFINITE
    pos[7, 0] = np.inf
    vel[123, 1] = 0.0
    d = hhs.DeviceParticles.from_host(pos, vel, m)
    with pytest.raises(hhs.HHSError) as ei:
        hhs.energy(d.pos, d.vel, d.m, 1e-8)
    assert ei.value.code == hhs.HHS_ERR_NONFINITE
