import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("pppm-tdbem_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def scenes_mod():
    import scenes
    return scenes


@pytest.fixture(scope="session")
def amd():
    import pppm_amd
    pppm_amd.load_library()
    return pppm_amd


def random_small_triangles(rng, n, res, h, size=0.1, margin=4):
    """random_pppm / regular_random_pppm of tests/case_generator.h:25-79 (seeded): small
    axis-aligned triangles scattered over the grid interior."""
    verts = np.zeros((3 * n, 3), np.float32)
    for i in range(n):
        v0 = np.array([0, 0, 1.0]) * rng.random() * rng.choice([-1, 1])
        v1 = np.array([1.0, 0, 0]) * rng.random() * rng.choice([-1, 1])
        v2 = np.array([0, 1.0, 0]) * rng.random() * rng.choice([-1, 1])
        off = (rng.random(3) * (res - 2 * margin) + margin) * h
        verts[3 * i:3 * i + 3] = np.stack([v0, v1, v2]) * size * h + off
    tris = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return verts, tris
