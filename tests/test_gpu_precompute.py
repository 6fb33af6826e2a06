"""Precompute-path API (s2fft.precompute_transforms) with method="hip" (needs a GPU).

The reference's precompute transforms compute the same spin-spherical and
Wigner transforms as its on-the-fly path, from a stored kernel; here they are
checked against the oracle at the FP64 tolerance of the plain kernels
(reference tests/test_spherical_precompute.py:57-58 use 1e-12).
"""

import numpy as np
import pytest

from oracle import signals, transforms as otr
from tests.conftest import max_rel

from s2fft_amd.precompute_transforms import construct, spherical as pspin

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.mark.parametrize("sampling,nside,L,spin,reality", [
    ("mw", None, 12, 0, True), ("mwss", None, 10, 2, False), ("gl", None, 9, -1, False),
    ("dh", None, 8, 0, False), ("healpix", 8, 16, 0, True)])
def test_precompute_spherical_matches_oracle(rng, sampling, nside, L, spin, reality):
    flm = signals.generate_flm(rng, L, 0, spin, reality)
    ker_i = construct.spin_spherical_kernel(L, spin, reality, sampling, nside, forward=False)
    f = pspin.inverse(flm, L, spin, ker_i, sampling, reality, "hip", nside)
    f_ref = otr.inverse(flm, L, spin, nside, sampling, reality)
    assert max_rel(f, f_ref) < TOL
    ker_f = construct.spin_spherical_kernel(L, spin, reality, sampling, nside, forward=True)
    fin = np.real(f_ref) if reality else f_ref
    got = pspin.forward(fin, L, spin, ker_f, sampling, reality, "hip", nside)
    ref = otr.forward(fin, L, spin, nside, sampling, reality)
    assert max_rel(got, ref) < TOL


def test_precompute_reality_spin_falls_back_to_complex(rng):
    L, spin = 8, 2
    flm = signals.generate_flm(rng, L, 0, spin, False)
    with pytest.warns(UserWarning):
        f = pspin.inverse(flm, L, spin, None, "mw", True, "hip")
    assert max_rel(f, otr.inverse(flm, L, spin, sampling="mw")) < TOL
