"""Seeded band-limited test signals with the reference's draw order.

Same distribution and NumPy generator call order as
s2fft/utils/signal_generator.py:10-123, so a seed gives the same flm here and
there: m = 0 real N(0,1); m > 0 complex with Re and Im each N(0,1), drawn in
np.tril_indices order; m < 0 by conjugate symmetry (reality) or fresh draws.
"""

import numpy as np


def complex_normal(rng, size, var=1.0):
    re = rng.standard_normal(size)
    im = rng.standard_normal(size)
    return np.asarray((re + 1j * im) * (var / 2) ** 0.5).astype(np.complex128)


def generate_flm(rng, L, L_lower=0, spin=0, reality=False):
    flm = np.zeros((L, 2 * L - 1), dtype=np.complex128)
    lo = max(L_lower, abs(spin))
    flm[lo:L, L - 1] = rng.standard_normal(L - lo)
    el, m = np.tril_indices(L, k=-1, m=L)
    m = m + 1
    keep = el >= lo
    el, m = el[keep], m[keep]
    flm[el, L - 1 + m] = complex_normal(rng, len(m), var=2)
    if reality:
        flm[el, L - 1 - m] = (-1.0) ** m * np.conj(flm[el, L - 1 + m])
    else:
        flm[el, L - 1 - m] = complex_normal(rng, len(m), var=2)
    return flm
