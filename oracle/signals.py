"""Seeded harmonic coefficients with the reference's draw order (oracle).

Test infrastructure only -- see ``oracle/__init__.py``.
"""

import numpy as np


def complex_normal(rng, size, var=1.0):
    """Re then Im standard normals scaled by sqrt(var/2).

    s2fft/utils/signal_generator.py:10-39 (both arrays drawn before combining).
    """
    re = rng.standard_normal(size)
    im = rng.standard_normal(size)
    return np.asarray((re + 1j * im) * (var / 2) ** 0.5).astype(np.complex128)


def _complex_positions(L, min_el):
    """(ell, m>0) positions in np.tril_indices order.  signal_generator.py:42-73."""
    el, m = np.tril_indices(L, k=-1, m=L)
    m = m + 1
    keep = el >= min_el
    return el[keep], m[keep]


def generate_flm(rng, L, L_lower=0, spin=0, reality=False):
    """s2fft/utils/signal_generator.py:76-123."""
    flm = np.zeros((L, 2 * L - 1), dtype=np.complex128)
    lo = max(L_lower, abs(spin))
    flm[lo:L, L - 1] = rng.standard_normal(L - lo)
    el, m = _complex_positions(L, lo)
    flm[el, L - 1 + m] = complex_normal(rng, len(m), var=2)
    if reality:
        flm[el, L - 1 - m] = (-1.0) ** m * np.conj(flm[el, L - 1 + m])
    else:
        flm[el, L - 1 - m] = complex_normal(rng, len(m), var=2)
    return flm


def generate_flmn(rng, L, N=1, L_lower=0, reality=False):
    """s2fft/utils/signal_generator.py:126-201."""
    flmn = np.zeros((2 * N - 1, L, 2 * L - 1), dtype=np.complex128)
    for n in range(-N + 1, N):
        lo = max(L_lower, abs(n))
        if reality:
            if n == 0:
                flmn[N - 1, lo:L, L - 1] = rng.standard_normal(L - lo)
            elif n > 0:
                flmn[N - 1 + n, lo:L, L - 1] = complex_normal(rng, L - lo, var=2)
                flmn[N - 1 - n, lo:L, L - 1] = (-1.0) ** n * np.conj(flmn[N - 1 + n, lo:L, L - 1])
        else:
            flmn[N - 1 + n, lo:L, L - 1] = complex_normal(rng, L - lo, var=2)
        el, m = _complex_positions(L, lo)
        flmn[N - 1 + n, el, L - 1 + m] = complex_normal(rng, len(m), var=2)
        if reality:
            flmn[N - 1 - n, el, L - 1 - m] = (-1.0) ** (m + n) * np.conj(flmn[N - 1 + n, el, L - 1 + m])
        else:
            flmn[N - 1 + n, el, L - 1 - m] = complex_normal(rng, len(m), var=2)
    return flmn
