"""Test configuration: the ``gpu`` marker and import paths.

``-m "not gpu"`` runs the oracle checks, host logic and C-ABI load/export
checks (no GPU needed); ``-m gpu`` runs the parity tests through the C-ABI.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "frequencydriftratetransforms.jl_amd"
for p in (str(ROOT), str(PKG)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libzdt_hip.so)")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import fdrt_hip
    fdrt_hip.lib()
    return torch.device("cuda", 0)
