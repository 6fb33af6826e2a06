"""Shared pytest configuration.

Markers: ``gpu`` tests need an MI355X (they run the HIP kernels through the C
ABI); everything else runs on the CPU.  The default seed mirrors the
reference's tests/conftest.py:15.
"""

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

DEFAULT_SEED = 8966433580120847635


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture
def rng():
    return np.random.default_rng(DEFAULT_SEED)


def max_rel(x, ref):
    """Norm-wise relative error ||x - ref||_inf / ||ref||_inf."""
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.max(np.abs(ref))
    return float(np.max(np.abs(x - ref)) / (den if den > 0 else 1.0))
