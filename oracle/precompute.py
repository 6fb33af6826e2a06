"""Stored Wigner-d tables and precompute kernels (oracle).

Test infrastructure only -- see ``oracle/__init__.py``.

* ``compute_all_slices`` -- s2fft/recursions/price_mcewen.py:377-643: the
  table dl[L-1+m, t, l] = d^l_{m,-spin}(beta_t), here assembled from the
  m-recursion restatement ``pm.dslices`` (whose values are pinned against
  Wigner's explicit sum in tests/test_oracle_pins.py).
* ``spin_spherical_kernel`` -- construct.py:16-161: K[t, l, m] =
  d^l_{m,-spin}(theta_t) sqrt((2l+1)/4pi) (x forward quadrature weights, x
  HEALPix ring phases), on the forward grid (MW/MWSS: MWSS at 2L with the
  poles) or the inverse grid.
* ``wigner_kernel`` -- construct.py:332-468: K[n, t, l, m] = d^l_{mn}(beta_t)
  x (2 pi / (2N-1) quadrature weights | (2l+1)/(8 pi^2)) x HEALPix phases.
* ``forward_transform`` / ``inverse_transform`` -- precompute_transforms/
  spherical.py:87-149, 314-380: the contractions with a kernel.
"""

import numpy as np

from . import ffts
from . import geometry as geo
from . import pm


def _zero_below(D, m, n, L0=0):
    """Zero the entries l < max(|m|, |n|): no Wigner-d value exists there; the
    reference's m-recursion leaves finite artefacts in them (its transforms
    never read those rows, its kernels are built for l >= |spin| only)."""
    lo = max(abs(m), abs(n)) - L0
    if lo > 0:
        D[:, :lo] = 0.0
    return D


def compute_all_slices(beta, L, spin):
    """dl (2L-1, len(beta), L); entries the m-recursion leaves NaN are 0 here
    (the reference's downstream sums drop them with nansum), and so are the
    entries l < max(|m|, |spin|)."""
    beta = np.asarray(beta, dtype=np.float64)
    dl = np.zeros((2 * L - 1, len(beta), L))
    for m, D in pm.dslices(beta, L, spin):
        dl[L - 1 + m] = _zero_below(np.nan_to_num(D), m, spin)
    return dl


def _dtable(beta, L, n, reality, exact=False):
    """[t, l, m - m0] = d^l_{m,n}(beta_t), m0 = 0 (reality) or -(L-1)."""
    from . import pins

    m0 = 0 if reality else -(L - 1)
    out = np.zeros((len(beta), L, L - m0))
    if exact:
        for m in range(m0, L):
            for l in range(max(abs(m), abs(n)), L):
                out[:, l, m - m0] = [pins.wigner_d(l, m, n, b) for b in beta]
        return out
    for m, D in pm.dslices(beta, L, -n, reality):
        out[:, :, m - m0] = _zero_below(np.nan_to_num(D), m, n)
    # pole rows (MWSS grids): the m-recursion divides by sin(beta); the
    # reference builds its kernels with Turok's l-slices (recursions/turok.py),
    # exact there -- the explicit sum gives those values
    for t in np.nonzero(np.abs(np.sin(beta)) < 1e-12)[0]:
        out[t] = 0.0
        for m in range(m0, L):
            for l in range(max(abs(m), abs(n)), L):
                out[t, l, m - m0] = pins.wigner_d(l, m, n, beta[t])
    return out


def _grid(L, sampling, nside, forward):
    s = sampling.lower()
    if forward and s in ("mw", "mwss"):
        return geo.thetas(2 * L, "mwss"), "mwss"
    return geo.thetas(L, s, nside), s


def spin_spherical_kernel(L, spin=0, reality=False, sampling="mw", nside=None, forward=True,
                          exact=False):
    """construct.py:16-161 (the Price-McEwen branch's values; every recursion
    gives the same d)."""
    if reality and spin != 0:
        reality = False
    th, s = _grid(L, sampling, nside, forward)
    dl = _dtable(th, L, -spin, reality, exact)
    dl = dl * np.sqrt((2 * np.arange(L) + 1) / (4 * np.pi))[None, :, None]
    if forward:
        dl = dl * geo.quad_weights_transform(L, s, nside)[:, None, None]
    if s == "healpix":
        m0 = 0 if reality else -(L - 1)
        m = np.arange(m0, L)
        phi0 = geo.healpix_phi0(np.arange(len(th)), nside)
        sign = -1.0 if forward else 1.0
        dl = dl * np.exp(sign * 1j * np.outer(phi0, m))[:, None, :]
    return dl


def wigner_kernel(L, N, reality=False, sampling="mw", nside=None, forward=False, exact=False):
    """construct.py:332-468."""
    th, s = _grid(L, sampling, nside, forward)
    n0 = 0 if reality else -(N - 1)
    planes = []
    for n in range(n0, N):
        dl = _dtable(th, L, n, False, exact)
        if forward:
            dl = dl * geo.quad_weights_transform(L, s, nside)[:, None, None] * (2 * np.pi / (2 * N - 1))
        else:
            dl = dl * ((2 * np.arange(L) + 1) / (8 * np.pi**2))[None, :, None]
        if s == "healpix":
            m = np.arange(-(L - 1), L)
            phi0 = geo.healpix_phi0(np.arange(len(th)), nside)
            dl = dl * np.exp((-1.0 if forward else 1.0) * 1j * np.outer(phi0, m))[:, None, :]
        planes.append(dl)
    return np.stack(planes)


def inverse_transform(flm, kernel, L, sampling, reality, spin, nside):
    """precompute_transforms/spherical.py:87-149."""
    s = sampling.lower()
    off = 1 if s in ("mwss", "healpix") else 0
    start = L - 1 if reality else 0
    ftm = np.zeros(geo.ftm_shape(L, s, nside), dtype=np.complex128)
    ftm[:, start + off :] = np.einsum("tlm,lm->tm", kernel, flm[:, start:])
    ftm *= (-1.0) ** spin
    if s == "healpix":
        if reality:
            ftm[:, off : start + off] = np.flip(np.conj(ftm[:, start + off + 1 :]), axis=-1)
        return ffts.healpix_ifft(ftm, L, nside, reality)
    if reality:
        return np.fft.irfft(ftm[:, start + off :], geo.nphi_equiang(L, s), axis=-1,
                            norm="forward")
    return np.fft.ifft(np.fft.ifftshift(ftm, axes=-1), axis=-1, norm="forward")


def forward_transform(f, kernel, L, sampling, reality, spin, nside):
    """precompute_transforms/spherical.py:314-380."""
    s = sampling.lower()
    if s == "mw":
        f = ffts.mw_to_mwss(f, L, spin)
    if s in ("mw", "mwss"):
        s = "mwss"
        f = ffts.upsample_by_two_mwss(f, L, spin)
    off = 1 if s in ("mwss", "healpix") else 0
    start = L - 1 if reality else 0
    if s == "healpix":
        ftm = ffts.healpix_fft(f, L, nside, reality)[:, off:]
        if reality:
            ftm = ftm[:, start:]
    elif reality:
        ftm = np.fft.rfft(np.real(f), axis=-1, norm="backward")
        if off:
            ftm = ftm[:, :-1]
    else:
        ftm = np.fft.fftshift(np.fft.fft(f, axis=-1, norm="backward"), axes=-1)[:, off:]
    flm = np.zeros((L, 2 * L - 1), dtype=np.complex128)
    flm[:, start:] = np.einsum("tlm,tm->lm", kernel, ftm)
    if reality:
        flm[:, :start] = np.flip((-1.0) ** (np.arange(1, L) % 2) * np.conj(flm[:, start + 1 :]),
                                 axis=-1)
    return flm * (-1.0) ** spin
