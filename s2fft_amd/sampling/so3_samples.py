"""SO(3) array shapes.  Same definitions as s2fft/sampling/so3_samples.py:6-102."""

from . import s2_samples


def f_shape(L, N, sampling="mw", nside=None):
    s = sampling.lower()
    if s in ("mw", "mwss", "dh", "gl"):
        return (2 * N - 1,) + s2_samples.f_shape(L, s)
    if s == "healpix":
        return 2 * N - 1, 12 * nside * nside
    raise ValueError(f"Sampling scheme sampling={sampling} not supported")


def flmn_shape(L, N):
    return 2 * N - 1, L, 2 * L - 1
