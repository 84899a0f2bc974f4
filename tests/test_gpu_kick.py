"""Kick parity: HIP kick (through the C ABI) vs the FP64 CPU oracle.

Force-error metric (tolerance 1e-6, from the 1e-6 relative bar for forces):
    err_i = |a_gpu,i - a_64,i| / sum_j m_j / (r_ij^2 + eps^2)
i.e. the error relative to the sum of the magnitudes of the pair terms.
A plain |da|/|a| is also reported, but a particle whose net force nearly
cancels can exceed 1e-6 there without any term being wrong.
"""
import numpy as np
import pytest

from conftest import force_scale, plummer

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _dev(hhs, pos, vel, m):
    return hhs.DeviceParticles.from_host(pos, vel, m)


def _check(a_gpu, a_ref, scale):
    err = np.linalg.norm(a_gpu - a_ref, axis=1) / scale
    return float(err.max())


# Sizes below ~32768 split the sources over several workgroups (split-K)
# with FP64 partials summed in chunk order; 5000 exercises a ragged split.
@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 257, 1000, 2048, 5000])
@pytest.mark.parametrize("eps2", [0.0, 1e-8])
def test_kick_self_matches_oracle(hhs, oracle, torch_cuda, n, eps2):
    pos, vel, m = plummer(hhs, n, seed=100 + n, eps2=eps2)
    d = _dev(hhs, pos, vel, m)
    dt, fac = 0.0078125, 0.5
    hhs.kick_self(d, dt=dt, fac=fac, eps2=eps2)
    torch_cuda.cuda.synchronize()
    vel_ref, acc_ref = oracle.kick(pos, vel, pos, m, dt, fac, eps2)
    acc = d.acc.cpu().numpy()
    if n == 1:
        assert np.all(acc == 0.0)
        return
    err = _check(acc, acc_ref, force_scale(pos, m, eps2))
    assert err < TOL, f"n={n} eps2={eps2}: max normalised force error {err:.3e}"
    dv = np.abs(d.vel.cpu().numpy() - vel_ref).max()
    assert dv <= TOL * dt * fac * force_scale(pos, m, eps2).max() + 1e-15


@pytest.mark.parametrize("s,w", [(1, 1), (1, 8), (2, 2), (2, 4), (4, 1), (4, 4), (4, 8)])
def test_kick_every_config(hhs, oracle, torch_cuda, s, w):
    n, eps2 = 777, 1e-8
    pos, vel, m = plummer(hhs, n, seed=7, eps2=eps2)
    d = _dev(hhs, pos, vel, m)
    hhs.kick_set_config(s, w)
    try:
        hhs.kick_self(d, dt=1.0, fac=1.0, eps2=eps2)
        torch_cuda.cuda.synchronize()
    finally:
        hhs.kick_set_config(0, 0)
    acc_ref = oracle.accel(pos, pos, m, eps2)
    err = _check(d.acc.cpu().numpy(), acc_ref, force_scale(pos, m, eps2))
    assert err < TOL, f"S={s} W={w}: {err:.3e}"


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("eps2", [0.0, 1e-8])
def test_kick_precision_modes(hhs, oracle, torch_cuda, precision, eps2):
    """MIXED32 (0), F64_RSQ32 (1) and F64 (2) all meet the force bar."""
    n = 3001
    pos, vel, m = plummer(hhs, n, seed=55, mass_model=1, eps2=eps2)
    d = _dev(hhs, pos, vel, m)
    hhs.kick_set_precision(precision)
    try:
        hhs.kick_self(d, dt=0.5, fac=1.0, eps2=eps2)
        torch_cuda.cuda.synchronize()
    finally:
        hhs.kick_set_precision(0)
    vel_ref, acc_ref = oracle.kick(pos, vel, pos, m, 0.5, 1.0, eps2)
    err = _check(d.acc.cpu().numpy(), acc_ref, force_scale(pos, m, eps2))
    assert err < TOL, f"precision {precision}: {err:.3e}"
    assert hhs.kick_get_precision() == 0


@pytest.mark.parametrize("n", [1, 2, 65, 1000, 3001])
@pytest.mark.parametrize("eps2", [0.0, 1e-8])
def test_kick_f64_exact_bit_exact(hhs, oracle, torch_cuda, n, eps2):
    """F64_EXACT: the oracle's operation order with IEEE sqrt/divide and the
    sources summed in index order -- acceleration and velocity bit for bit,
    including a coincident pair with eps2 = 0 (skipped, d2 == 0)."""
    pos, vel, m = plummer(hhs, n, seed=350 + n, mass_model=1, eps2=eps2)
    if n > 3:
        pos[3] = pos[2]
    d = _dev(hhs, pos, vel, m)
    hhs.kick_set_precision(hhs.KICK_F64_EXACT)
    try:
        hhs.kick_self(d, dt=0.0078125, fac=0.5, eps2=eps2)
        torch_cuda.cuda.synchronize()
    finally:
        hhs.kick_set_precision(0)
    vel_ref, acc_ref = oracle.kick(pos, vel, pos, m, 0.0078125, 0.5, eps2)
    assert d.acc.cpu().numpy().tobytes() == acc_ref.tobytes()
    assert d.vel.cpu().numpy().tobytes() == vel_ref.tobytes()


def test_kick_f64_exact_cross_bit_exact(hhs, oracle, torch_cuda):
    eps2 = 1e-8
    ps, vs, ms = plummer(hhs, 700, seed=21, eps2=eps2)
    pk, vk, mk = plummer(hhs, 333, seed=22, eps2=eps2)
    so, si = _dev(hhs, ps, vs, ms), _dev(hhs, pk, vk, mk)
    hhs.kick_set_precision(hhs.KICK_F64_EXACT)
    try:
        hhs.kick_cross(so, si, dt=0.3, fac=1.0, eps2=eps2)
        torch_cuda.cuda.synchronize()
    finally:
        hhs.kick_set_precision(0)
    vel_ref, acc_ref = oracle.kick(pk, vk, ps, ms, 0.3, 1.0, eps2)
    assert si.acc.cpu().numpy().tobytes() == acc_ref.tobytes()
    assert si.vel.cpu().numpy().tobytes() == vel_ref.tobytes()


@pytest.mark.parametrize("n_sink,n_src", [(1, 5), (300, 517), (517, 300), (64, 1), (129, 4096)])
def test_kick_cross_matches_oracle(hhs, oracle, torch_cuda, n_sink, n_src):
    eps2 = 1e-8
    ps, vs, ms = plummer(hhs, n_src, seed=11, eps2=eps2)
    pk, vk, mk = plummer(hhs, n_sink, seed=12, eps2=eps2)
    so, si = _dev(hhs, ps, vs, ms), _dev(hhs, pk, vk, mk)
    dt = 0.25
    hhs.kick_cross(so, si, dt=dt, fac=1.0, eps2=eps2)
    torch_cuda.cuda.synchronize()
    vel_ref, acc_ref = oracle.kick(pk, vk, ps, ms, dt, 1.0, eps2)
    scale = np.array([(ms / (((ps - p) ** 2).sum(1) + eps2)).sum() for p in pk])
    err = _check(si.acc.cpu().numpy(), acc_ref, scale)
    assert err < TOL, f"{n_sink}<-{n_src}: {err:.3e}"


def test_kick_zero_sources_leaves_velocity(hhs, torch_cuda):
    pos, vel, m = plummer(hhs, 100, seed=3)
    so, si = _dev(hhs, pos, vel, m), _dev(hhs, pos, vel, m)
    hhs.kick_cross(so, si, n_src=0, dt=1.0)
    torch_cuda.cuda.synchronize()
    assert np.array_equal(si.vel.cpu().numpy(), vel)
    assert np.all(si.acc.cpu().numpy() == 0.0)


def test_kick_two_body_closed_form(hhs, torch_cuda):
    # |a| = m / r^2 towards the other body
    pos = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, 0.0]])
    vel = np.zeros((2, 3))
    m = np.array([0.25, 0.75])
    d = _dev(hhs, pos, vel, m)
    hhs.kick_self(d, eps2=0.0)
    torch_cuda.cuda.synchronize()
    a = d.acc.cpu().numpy()
    r = 0.5
    u = (pos[1] - pos[0]) / r
    assert np.allclose(a[0], m[1] / r**2 * u, rtol=1e-6, atol=0)
    assert np.allclose(a[1], -m[0] / r**2 * u, rtol=1e-6, atol=0)


def test_kick_momentum_conserved(hhs, torch_cuda):
    n, eps2 = 4096, 1e-8
    pos, vel, m = plummer(hhs, n, seed=21, mass_model=1, eps2=eps2)
    d = _dev(hhs, pos, vel, m)
    hhs.kick_self(d, eps2=eps2)
    torch_cuda.cuda.synchronize()
    a = d.acc.cpu().numpy()
    f = (m[:, None] * a).sum(0)
    scale = (m[:, None] * np.abs(a)).sum(0)
    assert np.all(np.abs(f) <= 1e-6 * scale)


def test_kick_n16384_sampled(hhs, oracle, torch_cuda):
    """Plummer N=16384, eps=1e-4, against the FP64 oracle on 512 sinks."""
    n, eps2 = 16384, 1e-8
    pos, vel, m = plummer(hhs, n, seed=2, eps2=eps2)
    d = _dev(hhs, pos, vel, m)
    hhs.kick_self(d, eps2=eps2)
    torch_cuda.cuda.synchronize()
    idx = np.random.default_rng(0).choice(n, 512, replace=False)
    ref = oracle.accel_subset(pos, idx, pos, m, eps2)
    err = _check(d.acc.cpu().numpy()[idx], ref, force_scale(pos, m, eps2, idx))
    assert err < TOL, f"{err:.3e}"


# From 2^29 interactions per launch the source masses are converted to FP32
# in a separate pass before the force loop: large self and cross kicks, with
# close pairs (1e-7 apart) and far outliers (|x| ~ 1e3), on sampled sinks.
@pytest.mark.parametrize("n_sink,n_src,eps2", [(32768, 32768, 0.0), (32768, 32768, 1e-8),
                                               (4096, 131072, 1e-8)])
def test_kick_large_launch_sampled(hhs, oracle, torch_cuda, n_sink, n_src, eps2):
    rng = np.random.default_rng(n_sink + n_src)
    pos, vel, m = plummer(hhs, n_src, seed=7, eps2=eps2)
    pos[1] = pos[0] + 1e-7 * rng.standard_normal(3)
    pos[2] = pos[0] + np.array([1e-7, 0.0, 0.0])
    pos[3:6] *= 1e3
    src = _dev(hhs, pos, vel, m)
    if n_sink == n_src:
        hhs.kick_self(src, eps2=eps2)
        si, spos = src, pos
    else:
        spos = pos[:n_sink].copy()
        si = _dev(hhs, spos, vel[:n_sink].copy(), m[:n_sink].copy())
        hhs.kick_cross(src, si, eps2=eps2)
    torch_cuda.cuda.synchronize()
    idx = np.concatenate([np.arange(8), rng.choice(np.arange(8, n_sink), 504, replace=False)])
    ref = oracle.accel_subset(spos, idx, pos, m, eps2)
    # sink i sits at source i's position, so the self-kick scale applies
    err = _check(si.acc.cpu().numpy()[idx], ref, force_scale(pos, m, eps2, idx))
    assert err < TOL, f"{err:.3e}"


def test_kick_deterministic(hhs, torch_cuda):
    pos, vel, m = plummer(hhs, 3000, seed=9, eps2=1e-8)
    outs = []
    for _ in range(2):
        d = _dev(hhs, pos, vel, m)
        hhs.kick_self(d, eps2=1e-8)
        torch_cuda.cuda.synchronize()
        outs.append(d.acc.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


def test_kick_f32_storage(hhs, oracle, torch_cuda):
    import torch
    n, eps2 = 500, 1e-6
    pos, vel, m = plummer(hhs, n, seed=31, eps2=eps2)
    pos32 = pos.astype(np.float32).astype(np.float64)
    d = hhs.DeviceParticles.from_host(pos32, vel, m, dtype=torch.float32)
    hhs.kick_self(d, eps2=eps2)
    torch_cuda.cuda.synchronize()
    ref = oracle.accel(pos32, pos32, m.astype(np.float32).astype(np.float64), eps2)
    err = _check(d.acc.cpu().numpy().astype(np.float64), ref, force_scale(pos32, m, eps2))
    assert err < 1e-5  # FP32 storage: FP32 separations


def test_autotune_installs_a_valid_config(hhs, oracle, torch_cuda, tmp_path):
    path = tmp_path / "tune.txt"
    try:
        s, w, ms = hhs.autotune_kick(2048, 2, str(path))
        assert s in (1, 2, 4) and w in (1, 2, 4, 8) and ms > 0
        assert hhs.kick_get_config(2048, 2048) == (s, w)
        line = path.read_text().split()
        assert line[0] == "kick" and int(line[1]) == 11 and (int(line[2]), int(line[3])) == (s, w)
        pos, vel, m = plummer(hhs, 2048, seed=3, eps2=1e-8)
        d = _dev(hhs, pos, vel, m)
        hhs.kick_self(d, eps2=1e-8)
        torch_cuda.cuda.synchronize()
        err = _check(d.acc.cpu().numpy(), oracle.accel(pos, pos, m, 1e-8),
                     force_scale(pos, m, 1e-8))
        assert err < TOL
    finally:
        hhs.autotune_clear()
