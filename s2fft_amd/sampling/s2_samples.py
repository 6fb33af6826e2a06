"""Array shapes of the sampling schemes (host-side helpers of the product).

Same definitions as s2fft/sampling/s2_samples.py:4-155,446-492.
"""

SAMPLINGS = ("mw", "mwss", "dh", "gl", "healpix")


def ntheta(L=None, sampling="mw", nside=None):
    s = sampling.lower()
    if s != "healpix" and L is None:
        raise ValueError(f"Sampling scheme sampling={sampling} with L={L} not supported")
    if s in ("mw", "gl"):
        return L
    if s == "mwss":
        return L + 1
    if s == "dh":
        return 2 * L
    if s == "healpix":
        if nside is None:
            raise ValueError(f"Sampling scheme sampling={sampling} with nside={nside} not supported")
        return 4 * nside - 1
    raise ValueError(f"Sampling scheme sampling={sampling} not supported")


def nphi_equiang(L, sampling="mw"):
    s = sampling.lower()
    if s in ("mw", "gl", "dh"):
        return 2 * L - 1
    if s == "mwss":
        return 2 * L
    raise ValueError(f"Sampling scheme sampling={sampling} not supported")


def f_shape(L=None, sampling="mw", nside=None):
    s = sampling.lower()
    if s == "healpix":
        if nside is None:
            raise ValueError(f"Sampling scheme sampling={sampling} with nside={nside} not supported")
        return (12 * nside * nside,)
    if L is None:
        raise ValueError(f"Sampling scheme sampling={sampling} with L={L} not supported")
    return ntheta(L, s), nphi_equiang(L, s)


def flm_shape(L):
    return L, 2 * L - 1


def ftm_shape(L, sampling="mw", nside=None):
    ncol = 2 * L if sampling.lower() in ("mwss", "healpix") else 2 * L - 1
    return ntheta(L, sampling, nside), ncol
