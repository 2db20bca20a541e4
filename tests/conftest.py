"""pytest configuration: the `gpu` marker and shared helpers.

CPU tests (-m "not gpu") cover the oracle against the golden fixtures, the
host logic (maps, element kernels, meshes) and the C ABI's symbol table.
GPU tests (-m gpu) run the HIP path through the C ABI against the oracle.
"""
from __future__ import annotations

import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs HIP kernels (needs an MI355X)")


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible to torch (gpu tests need an MI355X)")
    return torch
