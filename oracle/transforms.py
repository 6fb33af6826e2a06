"""Forward / inverse spin-spherical harmonic and Wigner transforms (oracle).

Test infrastructure only -- see ``oracle/__init__.py``.
"""

import numpy as np

from . import ffts
from . import geometry as geo
from . import pm


def _norm(L, L_lower):
    return np.sqrt((2 * np.arange(L_lower, L) + 1) / (4 * np.pi))


def inverse(flm, L, spin=0, nside=None, sampling="mw", reality=False,
            precomps=None, L_lower=0):
    """Spin inverse transform.  s2fft/transforms/spherical.py:112-208."""
    s = sampling.lower()
    beta = geo.thetas(L, s, nside)
    off = 1 if s in ("mwss", "healpix") else 0
    flm = np.array(flm, dtype=np.complex128, copy=True)
    flm[L_lower:] *= _norm(L, L_lower)[:, None]
    ftm = pm.inverse_latitudinal_step(flm, beta, L, spin, nside, s, reality, precomps,
                                      L_lower)
    with np.errstate(invalid="ignore"):
        if s in ("mw", "mwss"):               # south pole, spherical.py:168-173
            ftm[-1] = 0
            ftm[-1, L - 1 + spin + off] = np.nansum(
                (-1.0) ** np.abs(np.arange(L_lower, L) - spin) * flm[L_lower:, L - 1 + spin])
        if s == "mwss":                       # north pole, 175-177
            ftm[0] = 0
            ftm[0, L - 1 - spin + off] = np.nansum(flm[L_lower:, L - 1 - spin])
    if s == "healpix":
        start = L - 1 if reality else 0
        ftm[:, start + off :] *= geo.ring_phase_shifts_hp(L, nside, False, reality)
    ftm *= (-1.0) ** spin
    if s == "healpix":
        if reality:
            ftm[:, off : L - 1 + off] = np.flip(np.conj(ftm[:, L + off :]), axis=-1)
        return ffts.healpix_ifft(ftm, L, nside, reality)
    if reality:
        return np.fft.irfft(ftm[:, L - 1 + off :], geo.nphi_equiang(L, s), axis=1,
                            norm="forward")
    return np.fft.ifft(np.fft.ifftshift(ftm, axes=1), axis=1, norm="forward")


def inverse_rings(flm, L, rings, spin=0, nside=None, sampling="mw", reality=False):
    """The inverse restricted to the rings ``rings``: their pixels, concatenated
    in the order given (complex128 like ``inverse`` for HEALPix, and for
    equiangular grids complex unless ``reality``).

    The steps of s2fft/transforms/spherical.py:112-208 for a subset of rings;
    every ring's latitudinal synthesis and azimuthal iDFT are independent of
    the others (otf_recursions.py:15-171 recurses per theta;
    healpix_ffts.py:335-371 folds and transforms per ring), so large grids
    stay cheap on the CPU.  The MW / MWSS pole rows (set in closed form by the
    reference) are not covered: rings must be interior rows there.
    """
    s = sampling.lower()
    rings = np.asarray(rings)
    if s == "mw":
        assert np.all(rings < L - 1), "MW row L-1 is the south pole"
    if s == "mwss":
        assert np.all((rings > 0) & (rings < L)), "MWSS rows 0 and L are the poles"
    beta = geo.thetas(L, s, nside)[rings]
    off = 1 if s in ("mwss", "healpix") else 0
    flm = np.array(flm, dtype=np.complex128, copy=True)
    flm *= _norm(L, 0)[:, None]
    ftm = pm.inverse_latitudinal_step(flm, beta, L, spin, nside, s, reality)
    if s == "healpix":
        start = L - 1 if reality else 0
        ftm[:, start + off :] *= geo.ring_phase_shifts_hp(L, nside, False, reality)[rings]
    ftm *= (-1.0) ** spin
    out = []
    if s != "healpix":
        nphi = geo.nphi_equiang(L, s)
        for k in range(len(rings)):
            if reality:
                out.append(np.fft.irfft(ftm[k, L - 1 + off :], nphi, norm="forward")
                           .astype(np.complex128))
            else:
                out.append(np.fft.ifft(np.fft.ifftshift(ftm[k]), norm="forward"))
        return np.concatenate(out)
    if reality:
        ftm[:, off : L - 1 + off] = np.flip(np.conj(ftm[:, L + off :]), axis=-1)
    for k, t in enumerate(rings):
        n = geo.nphi_ring(int(t), nside)
        row = ftm[k] if n == 2 * L else ffts.folding_row(ftm[k], n, L)
        if reality and n == 2 * L:
            out.append(np.fft.irfft(row[n // 2 :], n, norm="forward").astype(np.complex128))
        else:
            out.append(np.fft.ifft(np.fft.ifftshift(row), norm="forward"))
    return np.concatenate(out)


def forward(f, L, spin=0, nside=None, sampling="mw", reality=False, precomps=None,
            L_lower=0):
    """Spin forward transform.  s2fft/transforms/spherical.py:467-588."""
    s = sampling.lower()
    if s == "mw":
        f = ffts.mw_to_mwss(f, L, spin)
    if s in ("mw", "mwss"):
        s = "mwss"
        f = ffts.upsample_by_two_mwss(f, L, spin)
        beta = geo.thetas(2 * L, s)
    else:
        beta = geo.thetas(L, s, nside)
    w = geo.quad_weights_transform(L, s, nside)
    off = 1 if s in ("mwss", "healpix") else 0
    start = L - 1 if reality else 0

    if s == "healpix":
        ftm = ffts.healpix_fft(f, L, nside, reality)
    elif reality:
        t = np.fft.rfft(np.real(f), axis=1, norm="backward")
        if off:
            t = t[:, :-1]
        ftm = np.zeros(f.shape, dtype=np.complex128)
        ftm[:, L - 1 + off :] = t
    else:
        ftm = np.fft.fftshift(np.fft.fft(f, axis=1, norm="backward"), axes=1)
    ftm = ftm * w[:, None]
    if s == "healpix":
        ftm[:, start + off :] *= geo.ring_phase_shifts_hp(L, nside, True, reality)

    if s == "mwss":
        flm = pm.forward_latitudinal_step(ftm[1:-1], beta[1:-1], L, spin, nside, s,
                                          reality, precomps, L_lower)
        el = np.arange(L_lower, L)
        flm[L_lower:, L - 1 + spin] += (-1.0) ** np.abs(el - spin) * ftm[-1, L - 1 + spin + off]
        flm[L_lower:, L - 1 - spin] += ftm[0, L - 1 - spin + off]
    else:
        flm = pm.forward_latitudinal_step(ftm, beta, L, spin, nside, s, reality,
                                          precomps, L_lower)
    flm[L_lower:] *= _norm(L, L_lower)[:, None]
    if reality:                               # spherical.py:579-583
        msign = (-1.0) ** (np.arange(1, L) % 2)
        flm[:, :start] = np.flip(msign * np.conj(flm[:, start + 1 :]), axis=-1)
    flm[: max(abs(spin), L_lower)] = 0.0
    return flm * (-1.0) ** spin


def forward_iterative(f, L, spin=0, nside=None, sampling="mw", reality=False,
                      L_lower=0, n_iter=0, precomps=None):
    """flm = F(f); repeat n_iter: flm += F(f - I(flm)).

    s2fft/utils/iterative_refinement.py:9-46 as wired by
    s2fft/transforms/spherical.py:453-462.
    """
    F = lambda g: forward(g, L, spin, nside, sampling, reality, precomps, L_lower)
    flm = F(f)
    for _ in range(n_iter):
        flm = flm + F(f - inverse(flm, L, spin, nside, sampling, reality, None, L_lower))
    return flm


# ------------------------------------------------------------------ Wigner
def wigner_inverse(flmn, L, N, nside=None, sampling="mw", reality=False, L_lower=0):
    """SO(3) inverse.  s2fft/transforms/wigner.py:110-196."""
    s = sampling.lower()
    fban = np.zeros(geo.so3_f_shape(L, N, s, nside), dtype=np.complex128)
    flmn = np.array(flmn, dtype=np.complex128, copy=True)
    flmn[:, L_lower:] *= np.sqrt((2 * np.arange(L_lower, L) + 1) / (16 * np.pi**3))[:, None]
    n0 = 0 if reality else -N + 1
    for n in range(n0, N):
        fban[N - 1 + n] = (-1.0) ** n * inverse(flmn[N - 1 + n], L, -n, nside, s,
                                                reality if n == 0 else False, None, L_lower)
    ax = -2 if s == "healpix" else -3
    if reality:
        return np.fft.irfft(fban[N - 1 :], 2 * N - 1, axis=ax, norm="forward")
    return np.fft.ifft(np.fft.ifftshift(fban, axes=ax), axis=ax, norm="forward")


def wigner_forward(f, L, N, nside=None, sampling="mw", reality=False, L_lower=0):
    """SO(3) forward.  s2fft/transforms/wigner.py:442-533."""
    s = sampling.lower()
    flmn = np.zeros(geo.flmn_shape(L, N), dtype=np.complex128)
    ax = -2 if s == "healpix" else -3
    if reality:
        fban = np.fft.rfft(np.real(f), axis=ax, norm="backward")
    else:
        fban = np.fft.fftshift(np.fft.fft(f, axis=ax, norm="backward"), axes=ax)
    fban = fban * (2 * np.pi / (2 * N - 1))
    sgn = (-1.0) ** np.abs(np.arange(-L + 1, L))
    n0 = 0 if reality else -N + 1
    for n in range(n0, N):
        flmn[N - 1 + n] = (-1.0) ** n * forward(fban[n - n0], L, -n, nside, s,
                                                reality if n == 0 else False, None, L_lower)
        if reality and n != 0:
            flmn[N - 1 - n] = np.conj(np.flip(flmn[N - 1 + n] * sgn * (-1.0) ** n, axis=-1))
    flmn[:, L_lower:] *= np.sqrt(4 * np.pi / (2 * np.arange(L_lower, L) + 1))[:, None]
    return flmn
