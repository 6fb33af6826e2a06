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


def ld_ring_pixels(flm, L, nside, rings):
    """Pixels of HEALPix rings `rings` of the real spin-0 inverse of flm, with
    the latitudinal synthesis in 80-bit arithmetic (oracle.pins) and the
    azimuthal sum f(phi_p) = Re c_0 + 2 Re sum_{m>=1} c_m e^{i m phi_p} in
    double (error ~1e-16 of max |f|)."""
    from oracle import geometry, pins

    rings = list(rings)
    beta = geometry.thetas(L, "healpix", nside)[rings]
    flm_n = flm * np.sqrt((2 * np.arange(L) + 1) / (4 * np.pi))[:, None]
    ftm = pins.lat_inverse_longdouble(flm_n, beta, L)
    out = []
    m = np.arange(L)
    for k, t in enumerate(rings):
        phi = geometry.healpix_ring_phis(t, nside)
        vals = np.empty(len(phi))
        for a in range(0, len(phi), 512):
            e = np.exp(1j * np.outer(phi[a:a + 512], m))
            vals[a:a + 512] = 2 * np.real(e @ ftm[k]) - np.real(ftm[k, 0])
        out.append(vals)
    return np.concatenate(out)
