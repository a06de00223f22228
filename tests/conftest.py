"""Shared pytest set-up: package/oracle paths, the ``gpu`` marker, fixtures."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_ROOT = os.path.join(ROOT, "learning-based-attention-kernel-regression_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG_ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

LIB = os.path.join(PKG_ROOT, "laker", "liblaker_hip.so")
if not os.path.exists(LIB):
    sys.path.insert(0, ROOT)
    import __graft_entry__  # noqa: E402

    __graft_entry__.build()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (runs the HIP kernels)")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)

    return load


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda")
