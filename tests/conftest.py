import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.build()
    return oracle


def mismatch_report(a: np.ndarray, b: np.ndarray) -> str:
    """Pixel count and maximum LSB difference of two u16 depth arrays."""
    a = np.asarray(a).astype(np.int64)
    b = np.asarray(b).astype(np.int64)
    diff = np.abs(a - b)
    return f"{int((diff != 0).sum())} values differ, max {int(diff.max()) if diff.size else 0} LSB"


@pytest.fixture(scope="session")
def report():
    return mismatch_report
