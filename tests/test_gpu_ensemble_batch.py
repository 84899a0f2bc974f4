"""The batched-ensemble configuration at full size: 4096 Plummer systems,
N log-uniform in [8, 2048] (equal masses per system, seeded), eta = 0.01,
eps = 1e-4, t = 1 in root steps of 1/16, FP32 timesteps with exact
decisions, one persistent launch -- the workload bench.py times.

* every system ends with status HHS_OK at t = 1;
* energy: the median |dE/E| stays at the integrator's level (below 1e-6)
  and at least 75 % of the systems meet 1e-6 (the misses are the
  integrator's at this eta: bench.py reruns a sample through the FP64-only
  HOLD and finds the same errors);
* the recursion is accounted: calls, interactions and timestep pairs are
  positive and the interactions match the per-system sums;
* alone == in the mix, at full scale: eight systems spread over the N range
  (not the largest), integrated as a batch of their own, give bit for bit
  the states and results they have inside the 4096-system batch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS, ETA = 1e-4, 0.01


def _workload(hhs, n_systems=4096):
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(8), np.log(2048), n_systems)).astype(int)
    systems = [hhs.ic_plummer(int(n), 4000 + k, 0, True, EPS * EPS) for k, n in enumerate(sizes)]
    return sizes, systems


def test_ensemble_4096_systems_to_t1(hhs, torch_cuda):
    sizes, systems = _workload(hhs)
    print(f"ensemble: 4096 systems, N in [{sizes.min()}, {sizes.max()}], t = 1", flush=True)
    ens = hhs.Ensemble(systems, 1.0, 1.0 / 16, ETA, EPS, fast_timesteps=hhs.TIMESTEP_FP32)
    ens.run()
    torch_cuda.cuda.synchronize()
    r = ens.results()
    assert (r["status"] == 0).all()
    assert np.all(r["t"] == 1.0)
    de = np.abs((r["e1"] - r["e0"]) / r["e0"])
    print(f"|dE/E| median {np.median(de):.3e}, fraction <= 1e-6 {np.mean(de <= 1e-6):.4f}")
    assert np.median(de) < 1e-6
    assert np.mean(de <= 1e-6) >= 0.75
    assert (r["calls"] > 0).all() and (r["interactions"] >= 0).all()
    assert (r["ts_pairs"] > 0).all()

    order = np.argsort(sizes)
    pick = order[np.linspace(0, len(order) - 200, 8).astype(int)]
    print(f"alone: systems {pick.tolist()} (N {sizes[pick].tolist()})", flush=True)
    alone = hhs.Ensemble([systems[k] for k in pick], 1.0, 1.0 / 16, ETA, EPS,
                         fast_timesteps=hhs.TIMESTEP_FP32)
    alone.run()
    torch_cuda.cuda.synchronize()
    ra = alone.results()
    for j, k in enumerate(pick):
        for key in ("calls", "interactions", "ts_pairs", "max_level", "e0", "e1", "status"):
            assert ra[j][key] == r[k][key], (int(k), key)
        pa, va = alone.state(j)
        pb, vb = ens.state(int(k))
        assert pa.tobytes() == pb.tobytes() and va.tobytes() == vb.tobytes(), int(k)
