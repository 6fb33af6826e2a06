"""Wigner (SO(3)) transforms with ``method="hip"``.

Drop-in for s2fft.transforms.wigner.forward / inverse
(s2fft/transforms/wigner.py:13-107 and 346-439): f has shape [gamma, beta,
alpha] (or [gamma, pixel] for HEALPix) and flmn (2N-1, L, 2L-1).  The
gamma DFT runs in a HIP kernel; each n-plane is a spin -n transform through
the native plan whose recursion folds in the Wigner normalisation and the
(-1)^n factor (wigner.py:179-195, 506-533).
"""

import numpy as np
import torch

from .. import _lib
from ..sampling import so3_samples
from . import spherical as sph

METHODS = ("hip",)


def _check(method, N):
    if method not in METHODS:
        raise ValueError(f"Method {method} not recognised.")
    if N >= 8:
        raise Warning("Recursive transform may provide lower precision beyond N ~ 8")


def _plan(sampling, L, nside, n, L_lower, reality, device):
    sign = -1 if n % 2 else 1
    return sph.get_plan(sampling, L, nside, -n, L_lower, reality, _lib.NORM_WIGNER, sign,
                        device=device)


def forward_hip(f, L, N, nside=None, sampling="mw", reality=False, L_lower=0):
    """One SO(3) signal on the device -> flmn (2N-1, L, 2L-1) complex128."""
    dev = f.device
    G = 2 * N - 1
    npix = f[0].numel()
    stream = sph._stream(dev)
    fban = torch.empty((G,) + tuple(f.shape[1:]), dtype=torch.complex128, device=dev)
    src = f.contiguous()
    if reality:
        src = src.real.contiguous() if src.is_complex() else src.to(torch.float64)
        mode = 2
    else:
        src = src.to(torch.complex128)
        mode = 0
    _lib.check(_lib.lib().s2h_gamma_dft(N, npix, mode, 2 * np.pi / G, src.data_ptr(),
                                        fban.data_ptr(), stream), "gamma dft")
    flmn = torch.zeros((G, L, 2 * L - 1), dtype=torch.complex128, device=dev)
    n0 = 0 if reality else -N + 1
    for n in range(n0, N):
        plan = _plan(sampling, L, nside, n, L_lower, reality and n == 0, dev)
        flmn[N - 1 + n] = sph.forward_hip(fban[N - 1 + n][None], L, -n, nside, sampling,
                                          reality and n == 0, L_lower, plan=plan)[0]
    if reality:
        _lib.check(_lib.lib().s2h_wigner_reality_fill(flmn.data_ptr(), L, N, stream), "fill")
    return flmn


def inverse_hip(flmn, L, N, nside=None, sampling="mw", reality=False, L_lower=0):
    """flmn (2N-1, L, 2L-1) on the device -> one SO(3) signal."""
    dev = flmn.device
    G = 2 * N - 1
    fshape = so3_samples.f_shape(L, N, sampling, nside)
    npix = int(np.prod(fshape[1:]))
    stream = sph._stream(dev)
    n0 = 0 if reality else -N + 1
    fban = torch.empty((N if reality else G,) + tuple(fshape[1:]), dtype=torch.complex128,
                       device=dev)
    flmn = flmn.contiguous()
    for n in range(n0, N):
        plan = _plan(sampling, L, nside, n, L_lower, False, dev)
        g = sph.inverse_hip(flmn[N - 1 + n][None], L, -n, nside, sampling, False, L_lower,
                            plan=plan)[0]
        fban[n - n0] = g
    if reality:
        f = torch.empty(fshape, dtype=torch.float64, device=dev)
        mode = 3
    else:
        f = torch.empty(fshape, dtype=torch.complex128, device=dev)
        mode = 1
    _lib.check(_lib.lib().s2h_gamma_dft(N, npix, mode, 1.0, fban.data_ptr(), f.data_ptr(),
                                        stream), "gamma dft")
    return f


def inverse(flmn, L, N, nside=None, sampling="mw", method="hip", reality=False,
            precomps=None, L_lower=0, _ssht_backend=1):
    """SO(3) inverse (s2fft/transforms/wigner.py:13-107)."""
    _check(method, N)
    t, restore = sph._as_device_tensor(flmn, complex_only=True)
    with torch.cuda.device(t.device):
        f = inverse_hip(t, L, N, nside, sampling, reality, L_lower)
    return restore(f)


def forward(f, L, N, nside=None, sampling="mw", method="hip", reality=False,
            precomps=None, L_lower=0, _ssht_backend=1):
    """SO(3) forward (s2fft/transforms/wigner.py:346-439)."""
    _check(method, N)
    t, restore = sph._as_device_tensor(f)
    with torch.cuda.device(t.device):
        flmn = forward_hip(t, L, N, nside, sampling, reality, L_lower)
    return restore(flmn)
