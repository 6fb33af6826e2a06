"""Precompute-path Wigner transforms with ``method="hip"``.

Drop-in for s2fft.precompute_transforms.wigner.inverse / forward
(wigner.py:13-77 and 224-288).  ``kernel`` accepts the handle from
:func:`s2fft_amd.precompute_transforms.construct.wigner_kernel` (or a
reference kernel array, not used: the device regenerates the d-values).
"""

from ..transforms import wigner as _w

METHODS = ("hip",)


def _check(method):
    if method not in METHODS:
        raise ValueError(f"Method {method} not recognised.")


def inverse(flmn, L, N, kernel=None, sampling="mw", reality=False, method="hip", nside=None):
    """Inverse Wigner transform via the precompute API (wigner.py:13-77)."""
    _check(method)
    return _w.inverse(flmn, L, N, nside, sampling, "hip", reality)


def forward(f, L, N, kernel=None, sampling="mw", reality=False, method="hip", nside=None):
    """Forward Wigner transform via the precompute API (wigner.py:224-288)."""
    _check(method)
    return _w.forward(f, L, N, nside, sampling, "hip", reality)
