"""Pins the FP64 CPU oracle (CPU only).

The reference ships no golden vectors or known-answer tests (its tests
assert properties only), so every oracle function is pinned here by what
the mathematics fixes -- closed forms, invariants, brute force on tiny
inputs -- chosen so that a plausible slip (flipped sign, dropped softening,
swapped operands, off-by-one in the scan, lost stability) fails a check;
and, for the partition, by the reference's own cpu_split compiled from the
mounted reference tree (oracle/_ref).
"""
import math

import numpy as np
import pytest


# ------------------------------------------------------------------ kick
def test_accel_two_body_closed_form(oracle):
    pos = np.array([[0.1, 0.2, -0.3], [0.4, -0.2, -0.3]])
    m = np.array([2.0, 0.5])
    a = oracle.accel(pos, pos, m, 0.0)
    d = pos[1] - pos[0]
    r = np.linalg.norm(d)
    # attractive: body 0 is pulled towards body 1 with |a| = m1 / r^2
    assert np.allclose(a[0], m[1] * d / r**3, rtol=1e-14, atol=0)
    assert np.allclose(a[1], -m[0] * d / r**3, rtol=1e-14, atol=0)


def test_accel_softening_closed_form(oracle):
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.02]])
    m = np.array([1.0, 3.0])
    eps2 = 1e-4
    a = oracle.accel(pos, pos, m, eps2)
    r = 0.02
    assert a[0][2] == pytest.approx(3.0 * r / (r * r + eps2) ** 1.5, rel=1e-14)
    assert a[1][2] == pytest.approx(-1.0 * r / (r * r + eps2) ** 1.5, rel=1e-14)


def test_accel_cube_centre_is_zero(oracle):
    c = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], float) * 0.3
    a = oracle.accel(np.zeros((1, 3)), c, np.full(8, 0.125), 0.0)
    assert np.abs(a).max() < 1e-15


def test_accel_uses_source_masses_not_sink(oracle):
    si = np.array([[0.0, 0.0, 0.0]])
    so = np.array([[1.0, 0.0, 0.0], [0.0, -2.0, 0.0]])
    a = oracle.accel(si, so, np.array([0.5, 4.0]), 0.0)
    assert np.allclose(a[0], [0.5, -1.0, 0.0], rtol=1e-15)


def test_accel_momentum_and_bruteforce(oracle, hhs):
    pos, vel, m = hhs.ic_plummer(60, 5, 1, True, 0.0)
    a = oracle.accel(pos, pos, m, 1e-6)
    f = (m[:, None] * a).sum(0)
    assert np.abs(f).max() < 1e-13 * (m[:, None] * np.abs(a)).sum()
    # independent pairwise brute force, (i, j) double loop in Python
    b = np.zeros_like(a)
    for i in range(60):
        for j in range(60):
            d = pos[j] - pos[i]
            r2 = d @ d + 1e-6
            b[i] += m[j] * d / (r2 * math.sqrt(r2))
    assert np.allclose(a, b, rtol=1e-12, atol=1e-14)


def test_accel_skips_coincident_pairs(oracle):
    pos = np.array([[0.0, 0, 0], [0.0, 0, 0], [1.0, 0, 0]])
    a = oracle.accel(pos, pos, np.ones(3), 0.0)
    assert np.all(np.isfinite(a))
    assert a[0][0] == pytest.approx(1.0) and a[2][0] == pytest.approx(-2.0)


def test_kick_velocity_update(oracle):
    pos = np.array([[0.0, 0, 0], [2.0, 0, 0]])
    vel = np.array([[0.0, 1.0, 0], [0.0, -1.0, 0]])
    m = np.array([1.0, 1.0])
    v, a = oracle.kick(pos, vel, pos, m, dt=0.5, fac=0.25)
    assert np.allclose(a, [[0.25, 0, 0], [-0.25, 0, 0]])
    assert np.allclose(v, vel + a * 0.125)


# ------------------------------------------------------------- timesteps
def test_timestep_single_particle_is_huge(oracle):
    ts = oracle.timesteps(np.zeros((1, 3)), np.zeros((1, 3)), np.ones(1), 0.01)
    assert ts[0] == 1e30  # src/timestep.cuh:4 start value, j != i excluded


def test_timestep_circular_pair(oracle):
    r, M, eta = 0.5, 1.0, 0.03
    v = math.sqrt(M / r)
    pos = np.array([[-r / 2, 0, 0], [r / 2, 0, 0]])
    vel = np.array([[0, -v / 2, 0], [0, v / 2, 0]])
    ts = oracle.timesteps(pos, vel, np.array([0.5, 0.5]), eta)
    tff = eta / math.sqrt(2) * math.sqrt(r**3 / M)
    tfb = eta * r / v
    assert ts[0] == ts[1] == pytest.approx(min(tff, tfb), rel=1e-14)


def test_timestep_receding_and_approaching_pairs(oracle):
    # radial motion: u = (dr.dv)/r^2 = vr / r
    r, vr, M, eta = 1.0, 0.4, 1.0, 0.1
    for sgn in (+1, -1):
        pos = np.array([[0, 0, 0], [r, 0, 0]], float)
        vel = np.array([[0, 0, 0], [sgn * vr, 0, 0]], float)
        ts = oracle.timesteps(pos, vel, np.array([0.5, 0.5]), eta)
        u = sgn * vr / r
        tff = eta / math.sqrt(2) * r * math.sqrt(r / M)
        dff = min(1.5 * tff * u, 1.0)
        tff /= 1 - dff / 2
        tfb = eta * r / vr
        dfb = min(tfb * u * (1 + M / (vr * vr * r)), 1.0)
        tfb /= 1 - dfb / 2
        assert ts[0] == pytest.approx(min(tff, tfb), rel=1e-13)
        assert ts[0] == ts[1]


def test_timestep_cap_doubles_receding_flyby(oracle):
    # receding radial pair: the fly-by dtau is eta * (1 + M/(v^2 r)) >= 1 for
    # eta = 1.5, so it is capped at 1 and tau doubles (src/timestep.cuh:52-54)
    r, vr, M, eta = 1.0, 50.0, 1e-6, 1.5
    pos = np.array([[0, 0, 0], [r, 0, 0]], float)
    vel = np.array([[0, 0, 0], [vr, 0, 0]], float)
    ts = oracle.timesteps(pos, vel, np.array([M / 2, M / 2]), eta)
    assert ts[0] == pytest.approx(2 * eta * r / vr, rel=1e-13)


def test_timestep_min_over_partners_and_softening(oracle):
    pos = np.array([[0, 0, 0], [0.1, 0, 0], [0, 2.0, 0]], float)
    vel = np.zeros((3, 3))
    m = np.array([1.0, 1.0, 1.0])
    eta, eps2 = 0.05, 0.01
    ts = oracle.timesteps(pos, vel, m, eta, eps2)

    def tff(r2, mu):
        r = math.sqrt(r2)
        return eta / math.sqrt(2) * r * math.sqrt(r / mu)
    assert ts[0] == pytest.approx(tff(0.01 + eps2, 2.0), rel=1e-14)
    # body 2's nearer partner is body 0 (r^2 = 4), not body 1 (r^2 = 4.01)
    assert ts[2] == pytest.approx(tff(4.0 + eps2, 2.0), rel=1e-14)


# ------------------------------------------------------------ scan/split
@pytest.mark.parametrize("n", [0, 1, 2, 3, 512, 1023, 2151, 4090])
def test_exclusive_scan_is_prefix_sum(oracle, n):
    x = np.random.default_rng(n).integers(0, 5, n).astype(np.int32)
    out = oracle.exclusive_scan(x)
    assert np.array_equal(out, np.concatenate([[0], np.cumsum(x)[:-1]])[:n])


def _host(oracle, n, seed, ts):
    rng = np.random.default_rng(seed)
    return oracle.HostParticles(n, rng.standard_normal((n, 3)), rng.standard_normal((n, 3)),
                                rng.random(n), ts)


@pytest.mark.parametrize("n", [1, 2, 17, 1000])
def test_split_is_stable_partition(oracle, n):
    ts = np.random.default_rng(n + 1).random(n)
    h = _host(oracle, n, n, ts)
    slow, fast, ns, nf, pred, scan = oracle.split(h, 0.3)
    sel = ts <= 0.3
    if sel.sum() == 1:
        assert (ns, nf) == (n, 0)
        return
    assert nf == sel.sum() and ns == n - nf
    assert np.array_equal(pred, sel.astype(np.int32))
    assert np.array_equal(scan, np.concatenate([[0], np.cumsum(sel)[:-1]]).astype(np.int32))
    fi, si = np.nonzero(sel)[0], np.nonzero(~sel)[0]
    assert np.array_equal(fast.parent_id[:nf], fi) and np.array_equal(slow.parent_id[:ns], si)
    assert np.array_equal(fast.pos[:nf], h.pos[fi]) and np.array_equal(slow.vel[:ns], h.vel[si])
    assert np.array_equal(fast.id[:nf], h.id[fi])


def test_split_lone_fast_rule(oracle):
    n = 9
    ts = np.ones(n)
    ts[4] = 0.01
    h = _host(oracle, n, 3, ts)
    slow, fast, ns, nf, pred, _ = oracle.split(h, 0.5)
    assert (ns, nf) == (n, 0)
    assert pred.sum() == 1  # the raw predicate is still reported
    assert np.array_equal(slow.parent_id[:n], np.arange(n))
    assert np.array_equal(slow.pos[:n], h.pos)


def test_split_nan_is_slow(oracle):
    ts = np.array([0.1, np.nan, 0.2, np.nan])
    _, _, ns, nf, _, _ = oracle.split(_host(oracle, 4, 1, ts), 1.0)
    assert (ns, nf) == (2, 2)


REF_MATRIX = [1, 2, 3, 512, 1023, 1024, 2048, 2151, 4090, 524288, 588021]


@pytest.mark.parametrize("n", REF_MATRIX)
def test_cpu_split_matches_compiled_reference(oracle, n):
    """Restated cpu_split against the reference's own src/cpu_scan.h,
    compiled from /root/reference into oracle/_ref (values U[0,10),
    pivot 5, as in perf/perf_scan_multi_block.cu:186-244)."""
    if not oracle.ref_available():
        pytest.skip("reference tree not mounted: oracle/_ref not built")
    vals = np.random.default_rng(n).random(n) * 10.0
    for f32 in (False, True):
        v = vals.astype(np.float32) if f32 else vals
        r_arr, r_ns, r_nf = oracle.ref_cpu_split(v, 5.0, f32=f32)
        o_arr, o_ns, o_nf, rc = (oracle.cpu_split_f32(v, 5.0) if f32 else oracle.cpu_split(v, 5.0))
        assert rc == 0
        assert (o_ns, o_nf) == (r_ns, r_nf)
        assert np.array_equal(o_arr, r_arr)  # same unstable order, element for element
        # and the stable split has the same counts and multisets
        slow, fast = oracle.split_values(v.astype(np.float64), 5.0)
        assert (len(slow), len(fast)) == (r_ns, r_nf)
        assert np.array_equal(np.sort(fast), np.sort(r_arr[:r_nf].astype(np.float64)))


def test_cpu_split_empty(oracle):
    _, ns, nf, rc = oracle.cpu_split(np.zeros(0), 1.0)
    assert (ns, nf, rc) == (0, 0, 0)


# --------------------------------------------------------- drift/update
def test_drift_closed_form(oracle):
    pos = np.array([[1.0, 2.0, 3.0]])
    vel = np.array([[0.5, -0.25, 4.0]])
    assert np.array_equal(oracle.drift(pos, vel, 0.5, 0.5), pos + vel * 0.25)


def test_update_inverse_permutation(oracle):
    n = 50
    rng = np.random.default_rng(2)
    cur = oracle.HostParticles(n, rng.standard_normal((n, 3)), rng.standard_normal((n, 3)),
                               rng.random(n), rng.random(n))
    perm = rng.permutation(n).astype(np.int32)
    cur.parent_id[:] = perm
    prev = oracle.HostParticles(n)
    oracle.update(prev, cur)
    assert np.array_equal(prev.pos[perm], cur.pos) and np.array_equal(prev.m[perm], cur.m)
    assert np.array_equal(prev.timestep[perm], cur.timestep)


# ----------------------------------------------------------------- energy
def test_energy_two_body(oracle):
    pos = np.array([[0, 0, 0], [0.3, 0.4, 0]], float)
    vel = np.array([[1, 0, 0], [0, 2, 0]], float)
    m = np.array([2.0, 3.0])
    e = oracle.energy(pos, vel, m, 0.0)
    assert e == pytest.approx(0.5 * 2 * 1 + 0.5 * 3 * 4 - 6 / 0.5, rel=1e-15)
    e = oracle.energy(pos, vel, m, 0.11)
    assert e == pytest.approx(0.5 * 2 * 1 + 0.5 * 3 * 4 - 6 / math.sqrt(0.36), rel=1e-15)


# ------------------------------------------------------------------- HOLD
def _kepler(e=0.5):
    rp = 1.0 - e
    vp = math.sqrt((1 + e) / rp)
    pos = np.array([[-rp / 2, 0, 0], [rp / 2, 0, 0]])
    vel = np.array([[0, -vp / 2, 0], [0, vp / 2, 0]])
    return pos, vel, np.array([0.5, 0.5])


def test_hold_kepler_returns_after_one_period(oracle):
    pos, vel, m = _kepler()
    h = oracle.HostParticles(2, pos, vel, m)
    T = 2 * math.pi
    rc, st = oracle.hold(h, T, T / 4096, 0.01)
    assert rc == 0
    assert np.abs(h.pos - pos).max() < 1e-3
    assert st.simtime == pytest.approx(T, rel=1e-12)


def test_hold_second_order_energy_error(oracle, hhs):
    """|dE/E| over a fixed time scales as eta^2 for a hierarchical system."""
    pos, vel, m = hhs.ic_plummer(48, 11, 0, True, 1e-4)
    e0 = oracle.energy(pos, vel, m, 1e-4)
    errs = []
    for eta in (0.2, 0.1):
        h = oracle.HostParticles(48, pos, vel, m)
        rc, _ = oracle.hold(h, 0.25, 0.25, eta, 1e-4)
        assert rc == 0
        errs.append(abs(oracle.energy(h.pos, h.vel, h.m, 1e-4) / e0 - 1))
    assert 2.0 < errs[0] / errs[1] < 8.0


def test_hold_conserves_momentum_and_time(oracle, hhs):
    pos, vel, m = hhs.ic_plummer(64, 2, 1, True, 1e-6)
    h = oracle.HostParticles(64, pos, vel, m)
    rc, st = oracle.hold(h, 0.25, 0.125, 0.01, 1e-6)
    assert rc == 0
    p = (m[:, None] * h.vel).sum(0)
    assert np.abs(p).max() < 1e-13
    assert st.simtime == pytest.approx(0.25, rel=1e-14)  # leaf calls tile the interval
    assert st.max_level >= 2


def test_hold_depth_guard(oracle, hhs):
    pos, vel, m = hhs.ic_plummer(8, 1, 0, True, 0.0)
    pos[3] = pos[2]
    h = oracle.HostParticles(8, pos, vel, m)
    rc, _ = oracle.hold_step(h, 1.0 / 64, True, 0.01, 0.0, max_level=10)
    assert rc == 5  # ORACLE_ERR_DEPTH == HHS_ERR_DEPTH


# ------------------------------------------------ timestep corrections
# The criterion's corrections (src/timestep.cuh:42, :51) are first-order
# time derivatives of the raw timesteps along the pair's motion:
#   free-fall tau = eta/sqrt(2) sqrt(r^3/mu):  dtau/dt = 3 tau u / 2,
#   fly-by    tau = eta r / |v|:               dtau/dt = tau u (1 + mu/(v^2 r)),
# with u = (dr.dv)/r^2 and the relative acceleration -mu dr / r^3 (G = 1).
# The oracle applies tau <- tau / (1 - d/2) with d = dtau/dt (capped at 1),
# so d = 2 (1 - tau_raw / tau_oracle) can be read back from its output and
# compared with a central difference of the raw closed form along the true
# two-body trajectory (scipy's integrator, independent of the formula).  A
# wrong factor in either correction (3/2, the mu/(v^2 r) term, a sign)
# misses by O(1).


def _two_body_states(r0, v0, mu, times):
    from scipy.integrate import solve_ivp

    def f(_, y):
        r = y[:3]
        return np.concatenate([y[3:], -mu * r / np.linalg.norm(r) ** 3])

    sol = solve_ivp(f, (times[0], times[-1]), np.concatenate([r0, v0]), t_eval=times,
                    rtol=1e-13, atol=1e-15, method="DOP853")
    return sol.y[:3].T, sol.y[3:].T


def _pair_tau_and_d(oracle, r, v, m1, m2, eta, raw):
    """Oracle timestep of the pair (r = x_1 - x_2, v = v_1 - v_2) and the
    correction d read back from it against the raw closed form."""
    pos = np.array([0.5 * r, -0.5 * r])
    vel = np.array([0.5 * v, -0.5 * v])
    ts = oracle.timesteps(pos, vel, np.array([m1, m2]), eta, 0.0)
    assert ts[0] == ts[1]
    return ts[0], 2.0 * (1.0 - raw / ts[0])


@pytest.mark.parametrize("phase", [0.3, 1.2, 2.2, 3.0, 4.0, 5.5])
def test_timestep_free_fall_correction_is_time_derivative(oracle, phase):
    """Bound eccentric orbit (e = 0.6): free-fall is the smaller candidate;
    the oracle's correction equals d tau_ff / dt along the orbit."""
    m1, m2, eta = 0.7, 0.3, 1e-3
    mu = m1 + m2
    a, e = 1.0, 0.6
    rp = a * (1 - e)
    r0, v0 = np.array([rp, 0.0, 0.0]), np.array([0.0, np.sqrt(mu * (1 + e) / rp), 0.1 * 0.0])
    period = 2 * np.pi * np.sqrt(a ** 3 / mu)
    t, h = phase / (2 * np.pi) * period, 1e-4 * period
    R, V = _two_body_states(r0, v0, mu, np.array([0.0, t - h, t, t + h]))
    raw = [eta / np.sqrt(2) * np.sqrt(np.linalg.norm(x) ** 3 / mu) for x in R[1:]]
    fb = eta * np.linalg.norm(R[2]) / np.linalg.norm(V[2])
    assert raw[1] < 0.9 * fb  # the free-fall candidate wins
    tau, d = _pair_tau_and_d(oracle, R[2], V[2], m1, m2, eta, raw[1])
    d_true = (raw[2] - raw[0]) / (2 * h)
    assert abs(d_true) > 1e-5
    assert d == pytest.approx(d_true, rel=1e-6, abs=1e-12)


@pytest.mark.parametrize("t", [-3.0, -1.0, -0.2, 0.3, 1.5, 4.0])
def test_timestep_fly_by_correction_is_time_derivative(oracle, t):
    """Hyperbolic encounter (v_inf = 2, impact parameter 0.5): fly-by is the
    smaller candidate; the oracle's correction equals d tau_fb / dt along the
    trajectory, including the mu / (v^2 r) term from the gravitational
    deceleration."""
    m1, m2, eta = 0.6, 0.4, 1e-3
    mu = m1 + m2
    r0, v0 = np.array([-10.0, 0.5, 0.0]), np.array([2.0, 0.0, 0.0])
    t0 = 5.0  # closest approach near t0 + 0
    h = 1e-4
    R, V = _two_body_states(r0, v0, mu, np.array([0.0, t0 + t - h, t0 + t, t0 + t + h]))
    raw = [eta * np.linalg.norm(x) / np.linalg.norm(v) for x, v in zip(R[1:], V[1:])]
    ff = eta / np.sqrt(2) * np.sqrt(np.linalg.norm(R[2]) ** 3 / mu)
    assert raw[1] < 0.9 * ff  # the fly-by candidate wins
    tau, d = _pair_tau_and_d(oracle, R[2], V[2], m1, m2, eta, raw[1])
    d_true = (raw[2] - raw[0]) / (2 * h)
    assert abs(d_true) > 1e-6
    assert d == pytest.approx(d_true, rel=1e-6, abs=1e-12)
    # near the encounter the mu / (v^2 r) term is a sizeable part of d: a
    # correction without it would miss by far more than the tolerance
    r, v2 = np.linalg.norm(R[2]), np.dot(V[2], V[2])
    if r < 2.0:
        assert mu / (v2 * r) > 0.05


# ------------------------------------------------ calc_timestep flags
def test_hold_calc_timestep_flags_closed_form(oracle):
    """One root step of the triple (src/hold.cuh:74-178).  The root puts the
    pair in F (tau(0) <= dt) and body 3 in S.  The first recursion (dt/2)
    must reuse the root's timesteps (calc = false, :116-121): tau(0) > dt/2,
    a leaf.  The second (:168-173) must recompute them (calc = true) after
    the pair has moved for dt/2: tau(dt/2) <= dt/2, so it recurses; its own
    first child reuses tau(dt/2) > dt/4 (a leaf) and its second recomputes
    tau(3dt/4) > dt/4 (a leaf).  Calls per level [1, 2, 2] and 3*2 + 2 + 2 =
    10 timestep pairs.  The other flag choices give: swapped [1, 2] and 8;
    both recompute [1, 2, 2] and 14; neither [1, 2] and 6."""
    from conftest import flags_triple
    pos, vel, m, eta, dt, taus = flags_triple()
    assert dt / 2 < taus[0] <= dt and dt / 4 < taus[1] <= dt / 2 and taus[2] > dt / 4
    ts = oracle.timesteps(pos, vel, m, eta, 0.0)
    assert ts[0] == pytest.approx(taus[0], rel=1e-9) and ts[2] > 50 * dt
    host = oracle.HostParticles(3, pos, vel, m)
    rc, st = oracle.hold_step(host, dt, True, eta, 0.0)
    assert rc == 0
    assert [int(x) for x in st.calls_per_level[:st.max_level + 1]] == [1, 2, 2]
    assert st.ts_pairs == 10 and st.calls == 5
    assert st.simtime == dt
