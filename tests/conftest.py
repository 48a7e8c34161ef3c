import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (ROCm GPU) and the built HIP library")


@pytest.fixture(scope="session")
def small_scene():
    from mwdiem_amd import scene
    return scene.make_scene(60, (10.0, 10.0, 30.0), 11, name="small-60")


@pytest.fixture(scope="session")
def dense_scene(small_scene):
    """Neighbouring hulls overlap (about 1 contact in 12 candidate pairs)."""
    from mwdiem_amd import scene
    return scene.compact_dense(small_scene, gap=-0.55)


@pytest.fixture(scope="session")
def gpu():
    import torch
    from mwdiem_amd import _lib
    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    _lib.load()
    return torch.device("cuda", 0)


@pytest.fixture(scope="session")
def dense_phys(dense_scene):
    """The dense scene on the physical law (full stencil, midpoint contact point)."""
    from mwdiem_amd import scene
    sc = dense_scene.copy()
    sc.law_p = scene.Law4(mode=scene.PHYSICAL)
    sc.law_w = scene.Law4(mode=scene.PHYSICAL)
    return sc


@pytest.fixture(scope="session")
def impact_scene():
    """60 deposited hulls (touching, not overlapping) on the physical law."""
    from mwdiem_amd import scene
    return scene.make_scene(60, (6.0, 6.0, 16.0), 21, law=scene.Law4(mode=scene.PHYSICAL), name="impact-60")


@pytest.fixture(scope="session")
def large_scene():
    """The 100,000-hull settle scene (8-20 hull points, 200x200x400 mm box)."""
    from mwdiem_amd import scene
    return scene.config_large()
