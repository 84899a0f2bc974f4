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
