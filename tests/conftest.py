"""Shared pytest setup: the `gpu` marker, import paths, and small fixtures."""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libkf_hip.so")


@pytest.fixture(scope="session")
def ko():
    """The CPU oracle module (test infrastructure)."""
    from oracle import kf_oracle

    kf_oracle.lib()
    return kf_oracle


@pytest.fixture(scope="session")
def torch_gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU: run with -m 'not gpu' on CPU hosts")
    return torch


@pytest.fixture(scope="session")
def scene_room():
    """The seeded synthetic room (host renderer; no GPU needed)."""
    import kinectfusion_amd

    return kinectfusion_amd.SyntheticScene("room", seed=0)
