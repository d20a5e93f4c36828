"""Test setup: import paths, the ``gpu`` marker, shared configurations.

GPU tests are marked ``gpu`` and run on a ROCm GPU through the C-ABI library; everything else
runs on the CPU (the oracle, host physics, the library's exports and error paths).
"""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
for path in (ROOT / "ac-fdtd_amd", ROOT):
    if str(path) not in sys.path:
        sys.path.insert(0, str(path))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


def gpu_available() -> bool:
    import torch

    return torch.cuda.is_available()


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    with np.load(ROOT / "tests" / "golden" / "reference_vectors.npz") as data:
        return {key: data[key] for key in data.files}


@pytest.fixture
def require_gpu():
    if not gpu_available():
        pytest.skip("no ROCm GPU visible")
