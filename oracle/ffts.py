"""Longitudinal FFTs: HEALPix ragged rings and MW/MWSS resampling (oracle).

Test infrastructure only -- see ``oracle/__init__.py``.
"""

import numpy as np

from . import geometry as geo


def periodic_extension_row(fm, nphi, L):
    """Extend an fftshifted nphi-point spectrum to the 2L columns m in [-L, L).

    Column j of the result holds m = j - L and equals fm at bin m mod nphi
    (in fftshifted order).  s2fft/utils/healpix_ffts.py:85-117.
    """
    assert nphi <= 2 * L
    m = np.arange(-L, L)
    # fftshifted index of frequency k is k + nphi//2 for k in [-nphi//2, ...)
    k = (m + nphi // 2) % nphi
    return fm[k]


def folding_row(fm, nphi, L):
    """Alias the 2L columns m in [-L, L) onto nphi fftshifted bins.

    s2fft/utils/healpix_ffts.py:19-50: bins are indexed by m mod nphi with the
    same fftshift convention as ``periodic_extension_row``.
    """
    assert nphi <= 2 * L
    out = np.zeros(nphi, dtype=np.complex128)
    m = np.arange(-L, L)
    k = (m + nphi // 2) % nphi
    np.add.at(out, k, fm)
    return out


def ring_offsets(nside):
    """Start pixel of every HEALPix ring (ring-ordered map)."""
    nt = 4 * nside - 1
    n = np.array([geo.nphi_ring(t, nside) for t in range(nt)])
    return np.concatenate(([0], np.cumsum(n)[:-1])), n


def healpix_fft(f, L, nside, reality=False):
    """Per-ring DFT (unnormalised), fftshift, periodic extension to 2L columns.

    s2fft/utils/healpix_ffts.py:187-226.  With ``reality`` the rings whose
    nphi equals 2L use a real FFT and leave the m < 0 half at zero (213-215).
    """
    offs, nph = ring_offsets(nside)
    ftm = np.zeros(geo.ftm_shape(L, "healpix", nside), dtype=np.complex128)
    for t, (o, n) in enumerate(zip(offs, nph)):
        ring = f[o : o + n]
        if reality and n == 2 * L:
            row = np.zeros(n, dtype=np.complex128)
            row[n // 2 :] = np.fft.rfft(np.real(ring), norm="backward")[:-1]
            ftm[t] = row
        else:
            row = np.fft.fftshift(np.fft.fft(ring, norm="backward"))
            ftm[t] = row if n == 2 * L else periodic_extension_row(row, n, L)
    return ftm


def healpix_ifft(ftm, L, nside, reality=False):
    """Fold 2L columns onto each ring's bins, then unnormalised-forward iDFT.

    s2fft/utils/healpix_ffts.py:335-371.  The result is complex128 as in the
    reference, also for ``reality``.
    """
    assert L >= 2 * nside
    offs, nph = ring_offsets(nside)
    f = np.zeros(12 * nside * nside, dtype=np.complex128)
    for t, (o, n) in enumerate(zip(offs, nph)):
        row = ftm[t] if n == 2 * L else folding_row(ftm[t], n, L)
        if reality and n == 2 * L:
            f[o : o + n] = np.fft.irfft(row[n // 2 :], n, norm="forward")
        else:
            f[o : o + n] = np.fft.ifft(np.fft.ifftshift(row), norm="forward")
    return f


# ------------------------------------------------------------ resampling
def _mw_theta_extension_fourier(f_mw, L, spin):
    """MW grid -> 2L-1 theta rings in (theta, m) space.

    s2fft/utils/resampling.py:6-61 (periodic_extension with sampling="mw"):
    the phi-FFT of rings 1..L-1 is mirrored to rows L..2L-2 with the sign
    (-1)^(m + spin).  Returns the phi-spectrum (fftshifted, m in [-(L-1), L)).
    """
    B = f_mw.shape[0]
    ext = np.zeros((B, 2 * L - 1, 2 * L - 1), dtype=np.complex128)
    ext[:, :L, :] = f_mw
    ext = np.fft.fftshift(np.fft.fft(ext, axis=-1, norm="backward"), axes=-1)
    m = np.arange(-(L - 1), L)
    sign = (-1.0) ** np.abs(m) * (-1.0) ** abs(spin)
    ext[:, L : 2 * L - 1, :] = np.flip(ext[:, 0 : L - 1, :], axis=-2) * sign
    return ext


def mw_to_mwss(f_mw, L, spin=0):
    """MW (L, 2L-1) -> MWSS (L+1, 2L).  s2fft/utils/resampling.py:248-355."""
    squeeze = f_mw.ndim == 2
    f = f_mw[None] if squeeze else f_mw
    B = f.shape[0]
    # theta: periodic extension, shift by half a sample through Fourier space
    ext_fourier = _mw_theta_extension_fourier(f, L, spin)
    ext = np.fft.ifft(np.fft.ifftshift(ext_fourier, axes=-1), axis=-1, norm="backward")
    spec = np.zeros((B, 2 * L, 2 * L - 1), dtype=np.complex128)
    spec[:, 1:, :] = np.fft.fftshift(np.fft.fft(ext, axis=-2, norm="forward"), axes=-2)
    phase = np.exp(-1j * np.arange(-(L - 1), L) * np.pi / (2 * L - 1))
    spec[:, 1:, :] *= phase[None, :, None]
    g = np.fft.ifft(np.fft.ifftshift(spec, axes=-2), axis=-2, norm="forward")
    g = g[:, : L + 1, :]
    # phi: zero-pad the spectrum from 2L-1 to 2L samples (resampling.py:248-281)
    spec_phi = np.zeros((B, L + 1, 2 * L), dtype=np.complex128)
    spec_phi[:, :, 1:] = np.fft.fftshift(np.fft.fft(g, axis=-1, norm="forward"), axes=-1)
    out = np.fft.ifft(np.fft.ifftshift(spec_phi, axes=-1), axis=-1, norm="forward")
    return out[0] if squeeze else out


def upsample_by_two_mwss(f_mwss, L, spin=0):
    """MWSS (L+1, 2L) -> MWSS at 2L: (2L+1, 2L).

    s2fft/utils/resampling.py:64-160: spatial periodic extension over theta
    (rows L+1..2L-1 are the phi-rolled mirror of rows L-1..1 times (-1)^spin),
    theta-FFT, zero-pad to 4L rows, inverse FFT, keep rows 0..2L.
    """
    squeeze = f_mwss.ndim == 2
    f = f_mwss[None] if squeeze else f_mwss
    B = f.shape[0]
    ext = np.zeros((B, 2 * L, 2 * L), dtype=np.complex128)
    ext[:, : L + 1, :] = f
    mirror = np.flip(f[:, 1:L, :], axis=-2)
    ext[:, L + 1 :, :] = (-1.0) ** abs(spin) * np.fft.fftshift(mirror, axes=-1)
    spec = np.fft.fftshift(np.fft.fft(ext, axis=-2, norm="forward"), axes=-2)
    up = np.zeros((B, 4 * L, 2 * L), dtype=np.complex128)
    up[:, L : 3 * L, :] = spec
    out = np.fft.ifft(np.fft.ifftshift(up, axes=-2), axis=-2, norm="forward")
    out = out[:, : 2 * L + 1, :]
    return out[0] if squeeze else out
