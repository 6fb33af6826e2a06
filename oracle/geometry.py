"""Sampling geometry, quadrature weights and HEALPix ring phases (oracle).

Test infrastructure only -- see ``oracle/__init__.py``.
"""

import numpy as np

SAMPLINGS = ("mw", "mwss", "dh", "gl", "healpix")


def ntheta(L=None, sampling="mw", nside=None):
    """Number of theta rings.  s2fft/sampling/s2_samples.py:4-52."""
    s = sampling.lower()
    if s in ("mw", "gl"):
        return L
    if s == "mwss":
        return L + 1
    if s == "dh":
        return 2 * L
    if s == "healpix":
        return 4 * nside - 1
    raise ValueError(f"Sampling scheme sampling={sampling} not supported")


def nphi_equiang(L, sampling="mw"):
    """Number of phi samples of the equiangular grids.  s2_samples.py:85-123."""
    s = sampling.lower()
    if s in ("mw", "gl", "dh"):
        return 2 * L - 1
    if s == "mwss":
        return 2 * L
    raise ValueError(f"Sampling scheme sampling={sampling} not supported")


def ftm_shape(L, sampling="mw", nside=None):
    """(ntheta, number of m columns).  s2_samples.py:126-155."""
    ncol = 2 * L if sampling.lower() in ("mwss", "healpix") else 2 * L - 1
    return ntheta(L, sampling, nside), ncol


def f_shape(L=None, sampling="mw", nside=None):
    """Pixel array shape.  s2_samples.py:446-478."""
    if sampling.lower() == "healpix":
        return (12 * nside * nside,)
    return ntheta(L, sampling), nphi_equiang(L, sampling)


def flm_shape(L):
    """s2_samples.py:481-492: (L, 2L-1), column L-1+m."""
    return L, 2 * L - 1


def nphi_ring(t, nside):
    """Pixels on HEALPix ring t.  s2_samples.py:173-200."""
    if 0 <= t < nside - 1:
        return 4 * (t + 1)
    if nside - 1 <= t <= 3 * nside - 1:
        return 4 * nside
    if 3 * nside - 1 < t <= 4 * nside - 2:
        return 4 * (4 * nside - t - 1)
    raise ValueError(f"Ring t={t} not contained by nside={nside}")


def healpix_ring_z(t, nside):
    """cos(theta) of HEALPix ring(s) t.  s2_samples.py:284-308."""
    t = np.asarray(t, dtype=np.float64)
    z = np.empty_like(t)
    north = t < nside - 1
    south = t > 3 * nside - 1
    belt = ~north & ~south
    z[north] = 1.0 - (t[north] + 1.0) ** 2 / (3.0 * nside * nside)
    z[belt] = 4.0 / 3.0 - 2.0 * (t[belt] + 1.0) / (3.0 * nside)
    z[south] = (4.0 * nside - 1.0 - t[south]) ** 2 / (3.0 * nside * nside) - 1.0
    return z


def thetas(L=None, sampling="mw", nside=None):
    """Colatitudes of the rings.  s2_samples.py:203-281."""
    s = sampling.lower()
    if s == "gl":
        return np.flip(np.arccos(np.polynomial.legendre.leggauss(L)[0]))
    t = np.arange(ntheta(L, s, nside), dtype=np.float64)
    if s == "mw":
        return (2 * t + 1) * np.pi / (2 * L - 1)
    if s == "mwss":
        return 2 * t * np.pi / (2 * L)
    if s == "dh":
        return (2 * t + 1) * np.pi / (4 * L)
    if s == "healpix":
        return np.arccos(healpix_ring_z(t, nside))
    raise ValueError(f"Sampling scheme sampling={sampling} not supported")


def healpix_phi0(t, nside):
    """Azimuth of the first pixel of HEALPix ring(s) t.

    s2fft/utils/healpix_ffts.py:429-452 (p2phi_rings) and
    s2_samples.py:329-353 (p2phi_ring with p=0).
    """
    t = np.asarray(t)
    out = np.zeros(t.shape, dtype=np.float64)
    belt = (t + 1 >= nside) & (t + 1 <= 3 * nside)
    out = np.where(belt, 0.5 * ((t - nside + 2) % 2) * np.pi / (2 * nside), out)
    out = np.where(t + 1 > 3 * nside, 0.5 * np.pi / (2 * (4 * nside - t - 1)), out)
    out = np.where(t + 1 < nside, 0.5 * np.pi / (2 * (t + 1)), out)
    return out


def healpix_ring_phis(t, nside):
    """Pixel azimuths of ring t.  s2_samples.py:311-353."""
    n = nphi_ring(t, nside)
    p = np.arange(n, dtype=np.float64)
    if nside <= t + 1 <= 3 * nside:
        return np.pi / (2 * nside) * (p + 0.5 * ((t - nside + 2) % 2))
    if t + 1 > 3 * nside:
        return np.pi / (2 * (4 * nside - t - 1)) * (p + 0.5)
    return np.pi / (2 * (t + 1)) * (p + 0.5)


def ring_phase_shifts_hp(L, nside, forward=False, reality=False):
    """exp(-/+ i m phi0(t)) for m in [m_start, L).  healpix_ffts.py:482-509."""
    t = np.arange(ntheta(L, "healpix", nside))
    phi0 = healpix_phi0(t, nside)
    sign = -1.0 if forward else 1.0
    m = np.arange(0 if reality else -L + 1, L)
    return np.exp(sign * 1j * np.outer(phi0, m))


# ---------------------------------------------------------------- quadrature
def _mw_weight(m):
    """Fourier weights of the MW quadrature.  s2fft/utils/quadrature.py:291-320."""
    if m == 1:
        return 1j * np.pi / 2
    if m == -1:
        return -1j * np.pi / 2
    if m % 2 == 0:
        return 2.0 / (1.0 - m * m)
    return 0.0


def quad_weights_mwss_theta_only(L, spin=0):
    """quadrature.py:238-262."""
    w = np.zeros(2 * L, dtype=np.complex128)
    for i in range(-(L - 1) + 1, L + 1):
        w[i + L - 1] = _mw_weight(i - 1)
    wr = np.real(np.fft.fft(np.fft.ifftshift(w), norm="backward")) / (2 * L)
    q = wr[: L + 1].copy()
    q[1:L] = q[1:L] + (-1) ** spin * wr[-1:L:-1]
    return q


def quad_weights_mw_theta_only(L, spin=0):
    """quadrature.py:265-288."""
    w = np.array([_mw_weight(i) for i in range(-(L - 1), L)], dtype=np.complex128)
    w = w * np.exp(-1j * np.arange(-(L - 1), L) * np.pi / (2 * L - 1))
    wr = np.real(np.fft.fft(np.fft.ifftshift(w), norm="backward")) / (2 * L - 1)
    q = wr[:L].copy()
    q[: L - 1] = q[: L - 1] + (-1) ** spin * wr[-1 : L - 1 : -1]
    return q


def quad_weights_gl(L):
    """Gauss-Legendre weights x 2pi/(2L-1), Newton iteration on P_L.

    quadrature.py:125-161.
    """
    half = (L + 1) // 2
    i = np.arange(1, half + 1)
    z = np.cos(np.pi * (i - 0.25) / (L + 0.5))
    z_prev = np.full_like(z, 2.0)
    while np.max(np.abs(z - z_prev)) > 1e-14:
        p1 = np.ones_like(z)
        p2 = np.zeros_like(z)
        for j in range(1, L + 1):
            p3 = p2
            p2 = p1
            p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j
        pp = L * (z * p1 - p2) / (z * z - 1.0)
        z_prev = z
        z = z_prev - p1 / pp
    w = np.zeros(L, dtype=np.float64)
    w[i - 1] = 2.0 / ((1.0 - z**2) * pp * pp)
    w[L - i] = w[i - 1]
    return w * 2 * np.pi / (2 * L - 1)


def quad_weights_dh(L):
    """Driscoll-Healy weights.  quadrature.py:164-201."""
    th = thetas(L, "dh")
    k = np.arange(L)[:, None]
    w = np.sum(np.sin((2 * k + 1) * th[None, :]) / (2 * k + 1), axis=0)
    w = w * 2.0 / L * np.sin(th)
    return w * 2 * np.pi / (2 * L - 1)


def quad_weights_hp(nside):
    """4 pi / Npix on every ring.  quadrature.py:99-122."""
    return np.full(4 * nside - 1, 4 * np.pi / (12 * nside * nside))


def quad_weights_transform(L, sampling="mwss", nside=0):
    """Weights used by the forward transforms.  quadrature.py:7-50."""
    s = sampling.lower()
    if s == "mwss":
        return quad_weights_mwss_theta_only(2 * L, spin=0) * 2 * np.pi / (2 * L)
    if s == "dh":
        return quad_weights_dh(L)
    if s == "gl":
        return quad_weights_gl(L)
    if s == "healpix":
        return quad_weights_hp(nside)
    raise ValueError(f"Sampling scheme sampling={sampling} not supported")


def quad_weights(L=None, sampling="mw", spin=0, nside=None):
    """Integration weights on the native grid.  quadrature.py:53-96."""
    s = sampling.lower()
    if s == "mw":
        return quad_weights_mw_theta_only(L, spin) * 2 * np.pi / (2 * L - 1)
    if s == "mwss":
        return quad_weights_mwss_theta_only(L, spin) * 2 * np.pi / (2 * L)
    if s == "dh":
        return quad_weights_dh(L)
    if s == "gl":
        return quad_weights_gl(L)
    if s == "healpix":
        return quad_weights_hp(nside)
    raise ValueError(f"Sampling scheme sampling={sampling} not implemented")


# ---------------------------------------------------------------- SO(3)
def so3_f_shape(L, N, sampling="mw", nside=None):
    """s2fft/sampling/so3_samples.py:6-47: [gamma, beta, alpha] or [gamma, pix]."""
    if sampling.lower() == "healpix":
        return 2 * N - 1, 12 * nside * nside
    return (2 * N - 1,) + f_shape(L, sampling)


def flmn_shape(L, N):
    """so3_samples.py:50-65."""
    return 2 * N - 1, L, 2 * L - 1
