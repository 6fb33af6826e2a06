"""Precompute-path spin-spherical transforms with ``method="hip"``.

Drop-in for s2fft.precompute_transforms.spherical.inverse / forward
(spherical.py:19-84 and 226-311): same arguments and semantics, including the
fallback to the complex transform (with a warning) when ``reality`` is set
for a spin != 0 field and the iterative refinement of the forward.  The
``kernel`` argument accepts the handle returned by
:func:`s2fft_amd.precompute_transforms.construct.spin_spherical_kernel`; a
reference-style kernel array is accepted and not used (the device
regenerates the d-values, so its result is the transform the array would
encode for these L, spin, sampling, nside and reality).
"""

from warnings import warn

from ..transforms import spherical as sph
from .construct import KernelPlan

METHODS = ("hip",)


def _common(method, spin, reality):
    if method not in METHODS:
        raise ValueError(f"Method {method} not recognised.")
    if reality and spin != 0:
        warn("Reality acceleration only supports spin 0 fields. "
             "Defering to complex transform.", stacklevel=3)
        return False
    return reality


def _check_kernel(kernel, L, spin, sampling, nside, reality, forward):
    if isinstance(kernel, KernelPlan):
        want = (L, spin, reality, sampling.lower(), nside)
        got = (kernel.L, kernel.spin, kernel.reality, kernel.sampling.lower(), kernel.nside)
        if want != got:
            raise ValueError(f"kernel built for {got}, transform asked for {want}")


def inverse(flm, L, spin=0, kernel=None, sampling="mw", reality=False, method="hip",
            nside=None):
    """Inverse transform via the precompute API (spherical.py:19-84)."""
    reality = _common(method, spin, reality)
    _check_kernel(kernel, L, spin, sampling, nside, reality, False)
    return sph.inverse(flm, L, spin, nside, sampling, "hip", reality)


def forward(f, L, spin=0, kernel=None, sampling="mw", reality=False, method="hip", nside=None,
            iter=0):
    """Forward transform via the precompute API (spherical.py:226-311)."""
    reality = _common(method, spin, reality)
    _check_kernel(kernel, L, spin, sampling, nside, reality, True)
    return sph.forward(f, L, spin, nside, sampling, "hip", reality, iter=iter)
