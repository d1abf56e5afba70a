import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running parity case")


def cpu_threads() -> int:
    """Worker threads for the OpenMP oracle (the GPU machines grant 16 cores per GPU)."""
    return max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle

    return oracle.load()


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from ssmdlbm_amd import abi

    abi.load()
    return torch.device("cuda:0")
