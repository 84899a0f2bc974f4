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
