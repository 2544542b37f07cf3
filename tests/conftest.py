"""Shared fixtures.  Tests marked `gpu` need an MI355X; the rest run on the CPU."""
from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle  # noqa: the checker (oracle/oracle.py)

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def vrt():
    import voxelrt_amd

    voxelrt_amd.lib()
    return voxelrt_amd


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    return torch.device("cuda:0")
