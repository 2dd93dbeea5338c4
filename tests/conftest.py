"""Shared fixtures.  ``gpu`` marks tests that need an MI355X (ROCm GPU)."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libnan_hip.so")


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from nan_amd.device import Device

    return Device(0)


@pytest.fixture(scope="session")
def cornell():
    from nan_amd.mesh_colors import MeshColorsLayout
    from nan_amd.scene_arrays import flatten_scene
    from nan_amd.scenes import cornell_box

    node = cornell_box()
    layout = MeshColorsLayout.build(node)
    return node, layout, flatten_scene(node)
