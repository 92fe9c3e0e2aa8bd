import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libvoroffset_hip.so)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip_fp64():
    from voroffset_amd.morpho import VoronoiMorphoHip
    m = VoronoiMorphoHip("fp64")
    yield m
    m.close()


@pytest.fixture(scope="session")
def hip_fp32():
    from voroffset_amd.morpho import VoronoiMorphoHip
    m = VoronoiMorphoHip("fp32")
    yield m
    m.close()
