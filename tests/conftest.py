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
