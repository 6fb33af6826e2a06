"""Independent mathematics that pins the oracle (no reference code involved).

Test infrastructure only -- see ``oracle/__init__.py``.

* ``wigner_d``       -- Wigner's explicit sum for d^l_{mn}(beta), evaluated in
                        mpmath at 40 digits (standard convention
                        d^l_{mn}(beta) = <lm| exp(-i beta J_y) |ln>).
* ``sylm``           -- spin-weighted harmonics in the convention the transforms
                        implement: sY_lm = (-1)^s sqrt((2l+1)/4pi) e^{im phi}
                        d^l_{m,-s}(theta).
* ``ylm_scipy``      -- spin-0 harmonics from scipy.special.sph_harm_y
                        (Condon-Shortley phase), an independent library pin of
                        the spin-0 case of ``sylm``.
* ``healpix_pix2ang``-- the HEALPix ring-scheme closed form (Gorski et al. 2005,
                        eqs. 4-9), independent of the ring-index formulas.
* ``dft``            -- brute-force discrete Fourier sums.
* ``lat_inverse_longdouble`` -- the spin-0 latitudinal synthesis in 80-bit
                        long double (direct l-recursion from d^m_{m0}), a
                        high-precision pin for L up to a few hundred.
"""

import math

import mpmath
import numpy as np
import scipy.special as sps

mpmath.mp.dps = 40


def wigner_d(l, m, n, beta):
    """d^l_{mn}(beta) by Wigner's explicit formula in multiprecision."""
    if abs(m) > l or abs(n) > l:
        return 0.0
    b = mpmath.mpf(beta) / 2
    cb, sb = mpmath.cos(b), mpmath.sin(b)
    pref = mpmath.sqrt(math.factorial(l + m) * math.factorial(l - m)
                       * math.factorial(l + n) * math.factorial(l - n))
    total = mpmath.mpf(0)
    for k in range(max(0, n - m), min(l + n, l - m) + 1):
        den = (math.factorial(l + n - k) * math.factorial(k)
               * math.factorial(m - n + k) * math.factorial(l - m - k))
        term = (-1) ** (m - n + k) * pref / den
        total += term * cb ** (2 * l + n - m - 2 * k) * sb ** (m - n + 2 * k)
    return float(total)


def wigner_d_table(L, m, n, betas):
    """[t, l] table of d^l_{mn}(beta_t) for l in [0, L)."""
    out = np.zeros((len(betas), L))
    for l in range(L):
        for t, b in enumerate(betas):
            out[t, l] = wigner_d(l, m, n, b)
    return out


def sylm(l, m, s, theta, phi):
    """Spin-s harmonic sY_lm(theta, phi) = (-1)^s sqrt((2l+1)/4pi) e^{im phi} d^l_{m,-s}(theta)."""
    d = np.array([wigner_d(l, m, -s, th) for th in np.ravel(theta)]).reshape(np.shape(theta))
    return (-1.0) ** s * math.sqrt((2 * l + 1) / (4 * math.pi)) * d * np.exp(1j * m * np.asarray(phi))


def ylm_scipy(l, m, theta, phi):
    """scipy's Y_lm (polar angle theta, azimuth phi)."""
    return sps.sph_harm_y(l, m, theta, phi)


def synthesis_bruteforce(flm, L, spin, theta, phi, use_scipy=False):
    """f(theta_i, phi_i) = sum_{l,m} flm[l, L-1+m] sY_lm(theta_i, phi_i) at given points."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    out = np.zeros(theta.shape, dtype=np.complex128)
    for l in range(abs(spin), L):
        for m in range(-l, l + 1):
            c = flm[l, L - 1 + m]
            if c == 0:
                continue
            if use_scipy and spin == 0:
                out += c * ylm_scipy(l, m, theta, phi)
            else:
                out += c * sylm(l, m, spin, theta, phi)
    return out


def healpix_pix2ang(nside, p):
    """(theta, phi) of ring-ordered HEALPix pixel p (Gorski et al. 2005)."""
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    if p < ncap:                               # north polar cap
        i = int((1 + math.isqrt(1 + 2 * p)) // 2)
        while 2 * i * (i - 1) > p:
            i -= 1
        while 2 * (i + 1) * i <= p:
            i += 1
        j = p - 2 * i * (i - 1) + 1
        z = 1.0 - i * i / (3.0 * nside * nside)
        phi = math.pi / (2 * i) * (j - 0.5)
    elif p < npix - ncap:                      # equatorial belt
        q = p - ncap
        i = q // (4 * nside) + nside
        j = q % (4 * nside) + 1
        fodd = 1.0 if (i + nside) % 2 == 1 else 0.5
        z = 4.0 / 3.0 - 2.0 * i / (3.0 * nside)
        phi = math.pi / (2 * nside) * (j - fodd)
    else:                                      # south polar cap
        q = npix - p - 1
        i = int((1 + math.isqrt(1 + 2 * q)) // 2)
        while 2 * i * (i - 1) > q:
            i -= 1
        while 2 * (i + 1) * i <= q:
            i += 1
        j = 4 * i - (q - 2 * i * (i - 1))
        z = -1.0 + i * i / (3.0 * nside * nside)
        phi = math.pi / (2 * i) * (j - 0.5)
    return math.acos(z), phi


def dft(x):
    """X[k] = sum_p x[p] exp(-2 pi i k p / n)."""
    n = len(x)
    p = np.arange(n)
    return np.exp(-2j * np.pi * np.outer(p, p) / n) @ x


def lat_inverse_longdouble(flm, beta, L):
    """ftm[t, m] = sum_l flm[l, L-1+m] d^l_{m0}(beta_t), m = 0..L-1, in long double.

    The three-term recurrence in l of d^l_{m0} (Bonnet-type, normalised
    Wigner d) from d^m_{m0} = (-1)^m sqrt((2m)!/(m!^2)) (cos(b/2) sin(b/2))^m,
    every m at once (l outer): the start values are formed in log space, so
    the 80-bit exponent range (1e+-4932) holds them up to L of several
    thousand; values below it (the highest m at the most polar rings) are
    zero, and stay below it for every l < L.
    Returns complex128 of shape (len(beta), L).
    """
    ld = np.longdouble
    th = np.asarray(beta, dtype=np.float64).astype(ld)
    x = np.cos(th)
    cs = np.cos(th / 2) * np.sin(th / 2)
    nt = len(th)
    mv = np.arange(L).astype(ld)
    # log sqrt((2m)!/(m!^2)) by K_m = K_{m-1} (2m)(2m-1)/m^2
    inc = np.zeros(L, dtype=ld)
    inc[1:] = np.log((2 * mv[1:]) * (2 * mv[1:] - 1) / (mv[1:] * mv[1:]))
    half_logk = np.cumsum(inc) / 2
    sign = np.where(np.arange(L) % 2 == 1, ld(-1), ld(1))
    with np.errstate(under="ignore", divide="ignore"):
        d0 = sign[None, :] * np.exp(half_logk[None, :] + mv[None, :] * np.log(cs)[:, None])
    fr = np.asarray(flm[:, L - 1:].real, dtype=ld)
    fi = np.asarray(flm[:, L - 1:].imag, dtype=ld)
    dcur = np.zeros((nt, L), dtype=ld)
    dprev = np.zeros((nt, L), dtype=ld)
    accr = np.zeros((nt, L), dtype=ld)
    acci = np.zeros((nt, L), dtype=ld)
    for l in range(L):
        dcur[:, l] = d0[:, l]
        dprev[:, l] = 0
        k = l + 1
        accr[:, :k] += dcur[:, :k] * fr[l, :k]
        acci[:, :k] += dcur[:, :k] * fi[l, :k]
        if l == L - 1:
            break
        if l == 0:
            dn = x[:, None] * dcur[:, :1]
        else:
            m = mv[:k]
            D = ld(l) * np.sqrt((ld(l + 1) ** 2 - m * m) * ld(l + 1) ** 2)
            A = ld(2 * l + 1) * l * (l + 1) / D
            B = ld(l + 1) * np.sqrt((ld(l * l) - m * m) * ld(l * l)) / D
            dn = A * x[:, None] * dcur[:, :k] - B * dprev[:, :k]
        dprev[:, :k] = dcur[:, :k]
        dcur[:, :k] = dn
    return (accr.astype(np.float64) + 1j * acci.astype(np.float64)).astype(np.complex128)
