import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def golden():
    return GOLDEN


def make_field(shape, seed, kind="gauss", er=1e-3, n_gauss=4):
    """Small synthetic case (orig float64 (D,H,W), decp, xi)."""
    from extrem_z_amd import synth
    W, H, D = shape
    if kind == "gauss":
        o = synth.gaussians(shape, n_gauss, seed).astype(np.float64)
    elif kind == "noise":
        rng = np.random.default_rng(seed)
        o = rng.random((D, H, W)).astype(np.float32).astype(np.float64)
    elif kind == "grf":
        o = synth.grf(shape, seed).astype(np.float64)
    else:
        raise ValueError(kind)
    xi = synth.bound_from_er(o, er)
    return o, synth.quantise(o, xi), xi
