import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("parallel-hhs_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: long CPU oracle run")


@pytest.fixture(scope="session")
def hhs():
    import hhs as mod  # raises if libhhs.so is not built
    return mod


@pytest.fixture(scope="session")
def oracle():
    import oracle as mod
    return mod


@pytest.fixture(scope="session")
def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a visible GPU"
    return torch


def plummer(hhs_mod, n, seed, mass_model=0, eps2=0.0, virialize=True):
    return hhs_mod.ic_plummer(n, seed, mass_model, virialize, eps2)


def force_scale(pos, m, eps2, idx=None):
    """sum_j m_j / (r_ij^2 + eps2) for each sink (j != i, r > 0): the
    magnitude scale that the force-error metric is normalised by."""
    idx = np.arange(len(m)) if idx is None else np.asarray(idx)
    out = np.zeros(len(idx))
    for q, i in enumerate(idx):
        d = pos - pos[i]
        r2 = (d * d).sum(1) + eps2
        ok = r2 > 0
        out[q] = (m[ok] / r2[ok]).sum()
    return out


def flags_triple(eta=1.0, tau0=0.78, dt=1.0):
    """A hand-built triple whose HOLD call tree follows in closed form: bodies
    1 and 2 (mass 1e-10, so their motion over the step is a straight line to
    ~1e-10) approach head-on at relative speed v = 1; body 3 sits 50 away at
    rest.  Head-on, u = -v/r, so the fly-by correction is d = -eta and the
    pair's timestep is tau(t) = eta r(t) / (v (1 + eta/2)) with r(t) =
    r0 - v t: linear in time.  r0 is chosen so that tau(0) = tau0 = 0.78 dt.
    Then tau(dt/2) = 0.447 dt and tau(3 dt/4) = 0.280 dt, and body 3's
    timesteps are ~100 dt."""
    c = eta / (1.0 + eta / 2.0)
    r0 = tau0 * dt / c
    pos = np.array([[0.5 * r0, 0.0, 0.0], [-0.5 * r0, 0.0, 0.0], [0.0, 50.0, 0.0]])
    vel = np.array([[-0.5, 0.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.0, 0.0]])
    m = np.array([1e-10, 1e-10, 1e-10])
    taus = [c * (r0 - t) for t in (0.0, 0.5 * dt, 0.75 * dt)]
    return pos, vel, m, eta, dt, taus
