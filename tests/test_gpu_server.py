"""Continuous-batching server: a request's result is bit-identical alone and
in a mix (small systems through ensemble batches, large ones through the
per-system HOLD driver); admission respects the HBM budget; oversized
requests are refused."""
import numpy as np
import pytest

from conftest import plummer

pytestmark = pytest.mark.gpu

EPS, ETA = 1e-4, 0.01


def _reqs(hhs, sizes, seed0):
    rng = np.random.default_rng(seed0)
    out = []
    for k, n in enumerate(sizes):
        p, v, m = plummer(hhs, int(n), seed=seed0 + k, eps2=EPS * EPS)
        out.append((p, v, m, float(rng.uniform(0.02, 0.06)), 1.0 / 64))
    return out


def test_server_alone_equals_in_mix(hhs, torch_cuda):
    sizes = [8, 40, 130, 300, 700, 1500, 2600, 3100, 64, 900, 17, 2048]
    reqs = _reqs(hhs, sizes, 300)
    with hhs.Server(ETA, EPS, small_max_n=2048, max_n=4096, batch_streams=2,
                    big_workers=2) as srv:
        tickets = [srv.submit(p, v, m, te, dtr) for (p, v, m, te, dtr) in reqs]
        mixed = []
        for t in tickets:
            info = srv.wait(t)
            assert info["state"] == 3, info
            mixed.append((srv.fetch(t), info))
    for k in (0, 4, 6, 7, 11):
        p, v, m, te, dtr = reqs[k]
        with hhs.Server(ETA, EPS, small_max_n=2048, max_n=4096) as srv:
            t = srv.submit(p, v, m, te, dtr)
            info = srv.wait(t)
            (pa, va) = srv.fetch(t)
        (pm, vm), im = mixed[k]
        assert info["path"] == im["path"] == (0 if sizes[k] <= 2048 else 1)
        assert np.array_equal(pa, pm) and np.array_equal(va, vm), f"request {k}"
        assert info["result"] == im["result"]


def test_server_batch_matches_direct_ensemble(hhs, torch_cuda):
    reqs = _reqs(hhs, [50, 400], 77)
    with hhs.Server(ETA, EPS) as srv:
        ts = [srv.submit(p, v, m, te, dtr) for (p, v, m, te, dtr) in reqs]
        got = []
        for t in ts:
            srv.wait(t)
            got.append(srv.fetch(t))
    for k, (p, v, m, te, dtr) in enumerate(reqs):
        ens = hhs.Ensemble([(p, v, m)], te, dtr, ETA, EPS)
        ens.run()
        torch_cuda.cuda.synchronize()
        pe, ve = ens.state(0)
        assert np.array_equal(pe, got[k][0]) and np.array_equal(ve, got[k][1])


def test_server_budget_admission(hhs, torch_cuda):
    reqs = _reqs(hhs, [1000] * 6, 5)
    per = 1000 * (56 + 136) + 512
    budget = 2 * per + 100  # two small requests fit at a time
    with hhs.Server(ETA, EPS, hbm_budget_bytes=budget, small_max_n=2048) as srv:
        ts = [srv.submit(p, v, m, te, dtr) for (p, v, m, te, dtr) in reqs]
        infos = [srv.wait(t) for t in ts]
        st = srv.stats()
    assert all(i["state"] == 3 for i in infos)
    assert st["peak_bytes"] <= budget and st["used_bytes"] == 0
    # the third request could only be admitted after an earlier one finished
    admits = sorted(i["t_admit"] for i in infos)
    dones = sorted(i["t_done"] for i in infos)
    assert admits[2] >= dones[0]


def test_server_refuses_oversized_request(hhs, torch_cuda):
    p, v, m = plummer(hhs, 500, seed=1, eps2=EPS * EPS)
    with hhs.Server(ETA, EPS, hbm_budget_bytes=1000) as srv:
        with pytest.raises(hhs.HHSError) as ei:
            srv.submit(p, v, m, 0.1, 0.1)
    assert ei.value.code == hhs.HHS_ERR_BUDGET


def test_server_full_size_range_matches_direct_hold(hhs, oracle, torch_cuda):
    """The served workload's size range, N in [64, 65536] with one request
    above 32768, FP32 timesteps: every served result equals bit for bit the
    same system integrated on its own (a one-system ensemble up to
    small_max_n = 512, hhs_hold_f64 above), and a small and a mid-size
    request track the FP64 CPU HOLD oracle (same calls, positions within
    1e-6 of the rms)."""
    sizes = [64, 300, 700, 3000, 12000, 40000]
    t_ends = [0.5, 0.25, 0.125, 0.1, 0.1, 0.1]
    dtr, mode = 1.0 / 64, hhs.TIMESTEP_FP32
    reqs = [plummer(hhs, n, seed=5500 + k, eps2=EPS * EPS) for k, n in enumerate(sizes)]
    with hhs.Server(ETA, EPS, max_n=65536, max_level=40, fast_timesteps=mode) as srv:
        tickets = [srv.submit(p, v, m, te, dtr) for (p, v, m), te in zip(reqs, t_ends)]
        served = []
        for t in tickets:
            info = srv.wait(t)
            assert info["state"] == 3, info
            served.append((srv.fetch(t), info))
    for k, ((p, v, m), te) in enumerate(zip(reqs, t_ends)):
        (ps, vs), info = served[k]
        assert info["path"] == (0 if sizes[k] <= 512 else 1)
        if sizes[k] <= 512:
            ens = hhs.Ensemble([(p, v, m)], te, dtr, ETA, EPS, max_level=40, fast_timesteps=mode)
            ens.run()
            torch_cuda.cuda.synchronize()
            pd, vd = ens.state(0)
            calls = int(ens.results()[0]["calls"])
        else:
            d = hhs.DeviceParticles.from_host(p, v, m)
            g = hhs.Globals(len(m), ETA, EPS, max_level=40, fast_timesteps=mode)
            g.hold(d, te, dtr)
            torch_cuda.cuda.synchronize()
            pd, vd = d.pos.cpu().numpy(), d.vel.cpu().numpy()
            calls = g.stats()["calls"]
        assert pd.tobytes() == ps.tobytes() and vd.tobytes() == vs.tobytes(), sizes[k]
        assert calls == info["result"]["calls"]
    for k in (0, 2):
        (p, v, m), te = reqs[k], t_ends[k]
        host = oracle.HostParticles(len(m), p, v, m)
        rc, ost = oracle.hold(host, te, dtr, ETA, EPS * EPS)
        assert rc == 0
        (ps, vs), info = served[k]
        assert info["result"]["calls"] == ost.calls
        assert np.abs(ps - host.pos).max() <= 1e-6 * np.sqrt((host.pos**2).mean())
        assert np.abs(vs - host.vel).max() <= 1e-6 * np.sqrt((host.vel**2).mean())
