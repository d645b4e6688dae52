import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "spectrecon.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(scope="session")
def S():
    import spectrecon_amd

    return spectrecon_amd


@pytest.fixture(scope="session")
def O():
    import oracle

    return oracle


@pytest.fixture(scope="session")
def gpu(S):
    if S.device_count() == 0:
        pytest.skip("no HIP device")
    return S


def rand_sym_psfs(rng, px, pz, ny, nview):
    """Random point-symmetric PSFs normalised per plane (test/adjoint-project.jl:21-24)."""
    p = rng.random((px, pz, ny, nview)).astype(np.float32)
    p = p + p[::-1, ::-1]
    p = p / p.sum(axis=(0, 1), keepdims=True)
    # exact point symmetry after the division
    p = 0.5 * (p + p[::-1, ::-1])
    return np.asfortranarray(p.astype(np.float32))


def maxrel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def per_view_maxrel(g, r):
    """max over views of max|g-r|/max|r| within the view (the parity metric)."""
    return max(maxrel(g[:, :, k], r[:, :, k]) for k in range(r.shape[2]))
