"""Price-McEwen Wigner-d recursion and the latitudinal steps (oracle).

Test infrastructure only -- see ``oracle/__init__.py``.

The reference recurses over m at fixed (theta, ell): every ell starts on the
boundary m = +/-ell with a log-space value ``lrenorm`` and walks down to m = 0,
renormalising by the running magnitude after every step.  This module restates
that algorithm with NumPy, vectorised over (theta, ell) exactly like the
reference, and exposes it as a generator of d-slices so the inverse and the
forward steps share one restatement.
"""

import numpy as np

from . import geometry as geo


def precomputes(L, spin=0, sampling="mw", nside=None, forward=False, L_lower=0,
                betas=None):
    """Recursion coefficients with O(L^2) memory.

    s2fft/recursions/price_mcewen.py:118-255 (generate_precomputes_jax; the
    O(L^3) NumPy twin at 14-115 yields the same values).  Returns
    ``[lrenorm, vsign, cpi, cp2, indices]``.
    """
    mm = -spin
    L0 = L_lower
    if betas is None:
        if forward and sampling.lower() in ("mw", "mwss"):
            beta = geo.thetas(2 * L, "mwss")[1:-1]
        else:
            beta = geo.thetas(L, sampling, nside)
    else:
        beta = np.asarray(betas, dtype=np.float64)
    nt = len(beta)
    el = np.arange(L0, L)
    lt = np.log(np.abs(np.tan(-beta / 2.0)))
    c2 = np.cos(beta / 2.0)
    half = [el + mm + 1, el - mm + 1]

    with np.errstate(divide="ignore", invalid="ignore"):
        # cp(m) = cplus(l - m), price_mcewen.py:181-201
        cpi = np.zeros((L + 1, L - L0))
        cp2 = np.zeros((L + 1, L - L0))
        cpi[0] = 2.0 / np.sqrt(2 * el)
        for m in range(2, L + 1):
            cpi[m - 1] = 2.0 / np.sqrt(m * (2 * el + 1 - m))
            cp2[m - 1] = cpi[m - 1] / cpi[m - 2]
        for k in range(L0, L):
            cpi[:, k - L0] = np.roll(cpi[:, k - L0], L - k - 1)
            cp2[:, k - L0] = np.roll(cp2[:, k - L0], L - k - 1)

        # signs, price_mcewen.py:207-210
        msign = np.concatenate(((-1.0) ** np.arange(L - 1), np.ones(L)))
        lsign = (-1.0) ** np.abs(mm + el)
        vsign = np.outer(msign, lsign)
        vsign[: L - 1] *= (-1.0) ** abs(mm + 1 + L)

        # log boundary values by streaming log-ratio sums, price_mcewen.py:212-243
        lrenorm = np.zeros((2, nt, L - L0))
        row = np.outer(np.log(np.abs(c2)), 2.0 * el)
        i_all = np.arange(2 * L + 1)
        ratio = np.log((2 * el[None, :] + 2 - i_all[:, None]) / (i_all[:, None] - 1)) / 2
        for ind in range(2):
            lrenorm[ind] = np.where(half[ind][None, :] == 1, row, lrenorm[ind])
        for i in range(2, L + abs(mm) + 2):
            row = row + ratio[i][None, :] + lt[:, None]
            for ind in range(2):
                lrenorm[ind] = np.where(half[ind][None, :] == i, row, lrenorm[ind])

    indices = np.broadcast_to(el, (nt, L - L0)).copy()
    clean = lambda a: np.nan_to_num(a, nan=0.0, posinf=0.0, neginf=0.0)
    return [clean(lrenorm), vsign, clean(cpi), clean(cp2), indices]


def dslices(beta, L, spin, reality=False, L_lower=0, precomps=None,
            sampling="mw", nside=None, forward=False):
    """Yield ``(m, D)`` with D[t, ell - L_lower] = d^ell_{m,-spin}(beta_t).

    Follows the loop of s2fft/transforms/otf_recursions.py:259-356 (inverse)
    and 686-804 (forward): half i=0 covers m in [-(L-1), -1], half i=1 covers
    m in [0, L-1] (only i=1 when ``reality``).  Entries that the reference
    would drop through ``nansum`` are returned as NaN; callers sum with
    ``np.nansum``.
    """
    beta = np.asarray(beta, dtype=np.float64)
    mm = -spin
    L0 = L_lower
    nt = len(beta)
    el = np.arange(L0, L)
    c = np.cos(beta)
    s = np.sin(beta)
    omc = 1.0 - c
    half = [el + mm + 1, el - mm + 1]
    if precomps is None:
        precomps = precomputes(L, spin, sampling, nside, forward, L0, betas=beta)
    lrenorm, vsign, cpi, cp2, indices = precomps
    lrenorm = np.array(lrenorm, copy=True)
    lind = L - 1 - L0

    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        for i in range(2):
            if reality and i == 0:
                continue
            sind = 0 if i == 0 else 2 * L - 2      # vsign row of |m| = L-1
            sgn = 1 if i == 0 else -1
            col_m = lambda k: (-(L - 1) + k) if i == 0 else (L - 1 - k)

            it0 = np.ones((nt, L - L0))
            it1 = np.ones((nt, L - L0))
            lamb = (np.outer(omc, el + 1) + np.outer(c, 2 - L + el) - half[i][None, :]) / s[:, None]
            it1[:, lind:] = cpi[0, lind:] * it0[:, lind:] * lamb[:, lind:]

            # k = 0: |m| = L-1, only ell = L-1 (otf_recursions.py:285-293)
            D = np.zeros((nt, L - L0))
            D[:, lind:] = it0[:, lind:] * vsign[sind, lind:] * np.exp(lrenorm[i, :, lind:])
            yield col_m(0), D
            # k = 1: |m| = L-2, ell in {L-2, L-1} (296-304)
            if L - L0 >= 2:
                D = np.zeros((nt, L - L0))
                D[:, lind - 1 :] = (it1[:, lind - 1 :] * vsign[sind + sgn, lind - 1 :]
                                    * np.exp(lrenorm[i, :, lind - 1 :]))
                yield col_m(1), D

            entry = np.zeros((nt, L))
            for k in range(2, L - 1 + i):
                active = indices >= L - k - 1
                lamb = (np.outer(omc, el + 1) + np.outer(c, k - L + el + 1)
                        - half[i][None, :]) / s[:, None]
                entry[:, L0:] = np.where(active, cpi[k - 1] * it1 * lamb - cp2[k - 1] * it0,
                                         entry[:, L0:])
                entry[:, L - 1 - k] = 1.0
                D = entry[:, L0:] * vsign[sind + sgn * k] * np.exp(lrenorm[i])
                yield col_m(k), D
                mag = np.abs(entry[:, L0:])
                inv = 1.0 / mag
                it0 = np.where(active, inv * it1, it0)
                it1 = np.where(active, inv * entry[:, L0:], it1)
                lrenorm[i] = np.where(active, lrenorm[i] + np.log(mag), lrenorm[i])


def dslices_exact(beta, L, spin, reality=False, L_lower=0):
    """Yield ``(m, D)`` like ``dslices`` from Wigner's explicit sum (mpmath).

    For spin != 0 the reference's m-recursion meets exact zeros of
    intermediate values at some HEALPix colatitudes (e.g. cos(theta) = 1/6):
    its renormalisation 1/|d| (otf_recursions.py:164-169) turns them into NaN
    and nansum drops those terms.  This generator gives the mathematically
    exact d-values for those cases; it is slow and meant for L <= 16.
    """
    from . import pins

    beta = np.asarray(beta, dtype=np.float64)
    for m in range(0 if reality else -(L - 1), L):
        D = np.zeros((len(beta), L - L_lower))
        for l in range(max(abs(m), abs(spin), L_lower), L):
            D[:, l - L_lower] = [pins.wigner_d(l, m, -spin, b) for b in beta]
        yield m, D


_EXACT = {"on": False}


class exact_d:
    """Context manager: latitudinal steps use ``dslices_exact``."""

    def __enter__(self):
        _EXACT["on"] = True
        return self

    def __exit__(self, *exc):
        _EXACT["on"] = False


def _slices(beta, L, spin, reality, L_lower, precomps, sampling, nside, forward):
    if _EXACT["on"]:
        return dslices_exact(beta, L, spin, reality, L_lower)
    return dslices(beta, L, spin, reality, L_lower, precomps, sampling, nside, forward)


def _ftm_col_offset(sampling):
    return 1 if sampling.lower() in ("mwss", "healpix") else 0


def inverse_latitudinal_step(flm, beta, L, spin, nside=None, sampling="mw",
                             reality=False, precomps=None, L_lower=0):
    """ftm[t, m] = sum_ell flm[ell, m] d^ell_{m,-spin}(beta_t).

    s2fft/transforms/otf_recursions.py:15-171 / 174-410.  ``flm`` is (L, 2L-1)
    with column L-1+m; the result has the reference ftm layout for
    ``sampling`` (2L columns, column L+m, for mwss/healpix; else 2L-1).
    """
    off = _ftm_col_offset(sampling)
    ftm = np.zeros((len(beta), 2 * L - 1 + off), dtype=np.complex128)
    for m, D in _slices(beta, L, spin, reality, L_lower, precomps, sampling, nside, False):
        ftm[:, L - 1 + m + off] = np.nansum(D * flm[L_lower:, L - 1 + m][None, :], axis=1)
    return ftm


def forward_latitudinal_step(ftm, beta, L, spin, nside=None, sampling="mw",
                             reality=False, precomps=None, L_lower=0):
    """flm[ell, m] = sum_t ftm[t, m] d^ell_{m,-spin}(beta_t).

    s2fft/transforms/otf_recursions.py:432-588.
    """
    off = _ftm_col_offset(sampling)
    flm = np.zeros((L, 2 * L - 1), dtype=np.complex128)
    for m, D in _slices(beta, L, spin, reality, L_lower, precomps, sampling, nside, True):
        flm[L_lower:, L - 1 + m] = np.nansum(D * ftm[:, L - 1 + m + off][:, None], axis=0)
    return flm
