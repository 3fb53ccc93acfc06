import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "blade-volume_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def c1_scene():
    """16,384 sites U[-1,1]^3 (seed 0) with Qhull Delaunay adjacency, SH degree 0."""
    from bv import scenes

    return scenes.uniform_scene(16_384, seed=0)


@pytest.fixture(scope="session")
def small_scene():
    """512 uniform sites with Delaunay adjacency and SH degree 3 coefficients."""
    from bv import scenes

    rng = np.random.default_rng(11)
    n = 512
    xyz = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    off, nb = scenes.delaunay_csr(xyz)
    density = np.exp(rng.normal(size=n)).astype(np.float32)
    sh = rng.normal(scale=0.3, size=(n, 48)).astype(np.float32)
    return scenes.HostScene(xyz, density, sh, 3, off, nb)


def c1_camera():
    from bv import scenes

    return scenes.looking_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), np.deg2rad(60.0), 1.0, 100.0)
