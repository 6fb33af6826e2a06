"""Seeded band-limited test signals with the reference's draw order.

Same distribution and NumPy generator call order as
s2fft/utils/signal_generator.py:10-201, so a seed gives the same coefficients
here and there: m = 0 real N(0,1); m > 0 complex with Re and Im each N(0,1),
drawn in np.tril_indices order; m < 0 by conjugate symmetry (reality) or
fresh draws.  Wigner coefficients add the n planes in the order -N+1 .. N-1.
"""

import numpy as np


def complex_normal(rng, size, var=1.0):
    """Re then Im standard normals scaled by sqrt(var/2) (signal_generator.py:10-39)."""
    re = rng.standard_normal(size)
    im = rng.standard_normal(size)
    return np.asarray((re + 1j * im) * (var / 2) ** 0.5).astype(np.complex128)


def _positive_m(L, lo):
    """(l, m > 0) with l >= lo in np.tril_indices order (signal_generator.py:42-73)."""
    el, m = np.tril_indices(L, k=-1, m=L)
    m = m + 1
    keep = el >= lo
    return el[keep], m[keep]


def generate_flm(rng, L, L_lower=0, spin=0, reality=False):
    """(L, 2L-1) spin-spin coefficients (signal_generator.py:76-123)."""
    flm = np.zeros((L, 2 * L - 1), dtype=np.complex128)
    lo = max(L_lower, abs(spin))
    flm[lo:L, L - 1] = rng.standard_normal(L - lo)
    el, m = _positive_m(L, lo)
    flm[el, L - 1 + m] = complex_normal(rng, len(m), var=2)
    if reality:
        flm[el, L - 1 - m] = (-1.0) ** m * np.conj(flm[el, L - 1 + m])
    else:
        flm[el, L - 1 - m] = complex_normal(rng, len(m), var=2)
    return flm


def generate_flmn(rng, L, N=1, L_lower=0, reality=False):
    """(2N-1, L, 2L-1) Wigner coefficients (signal_generator.py:126-201).

    Per n (ascending): the m = 0 column (real for n = 0 of a real signal,
    complex for n > 0 with its n < 0 partner by symmetry, nothing for n < 0 of
    a real signal; complex for every n otherwise), then the m > 0 entries, then
    the m < 0 entries: for a real signal flmn[-n, l, -m] = (-1)^(m+n)
    conj(flmn[n, l, m]), else fresh draws.
    """
    flmn = np.zeros((2 * N - 1, L, 2 * L - 1), dtype=np.complex128)
    for n in range(-N + 1, N):
        k = N - 1 + n
        lo = max(L_lower, abs(n))
        if not reality:
            flmn[k, lo:L, L - 1] = complex_normal(rng, L - lo, var=2)
        elif n == 0:
            flmn[k, lo:L, L - 1] = rng.standard_normal(L - lo)
        elif n > 0:
            flmn[k, lo:L, L - 1] = complex_normal(rng, L - lo, var=2)
            flmn[N - 1 - n, lo:L, L - 1] = (-1.0) ** n * np.conj(flmn[k, lo:L, L - 1])
        el, m = _positive_m(L, lo)
        flmn[k, el, L - 1 + m] = complex_normal(rng, len(m), var=2)
        if reality:
            flmn[N - 1 - n, el, L - 1 - m] = (-1.0) ** (m + n) * np.conj(flmn[k, el, L - 1 + m])
        else:
            flmn[k, el, L - 1 - m] = complex_normal(rng, len(m), var=2)
    return flmn
