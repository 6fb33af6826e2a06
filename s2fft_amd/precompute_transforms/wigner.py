"""Precompute-path Wigner transforms with ``method="hip"``.

Drop-in for s2fft.precompute_transforms.wigner.inverse / forward
(wigner.py:13-77 and 224-288): the gamma DFT (``s2h_gamma_dft``), then per
n-plane the ring FFTs (+ MW/MWSS resampling at spin -n) and the stored
kernel's dense contraction (``s2h_precompute_forward`` / ``_inverse``,
einsum "ntlm,ntm->nlm" and "ntlm,nlm->ntm" of wigner.py:80-149, 290-380).
``kernel``: the tensor from :func:`.construct.wigner_kernel` or a
reference-made NumPy kernel of the same layout; None builds it on the device.
"""

import numpy as np
import torch

from .. import _lib
from ..sampling import so3_samples
from ..transforms import spherical as sph
from . import construct
from .spherical import _kernel_tensor

METHODS = ("hip",)


def _check(method):
    if method not in METHODS:
        raise ValueError(f"Method {method} not recognised.")


def _plane_plan(sampling, L, nside, n, device):
    return sph.get_plan(sampling, L, nside, -n, 0, False, _lib.NORM_WIGNER,
                        -1 if n % 2 else 1, device=device)


def _kernel(kernel, L, N, sampling, nside, reality, forward, device):
    if kernel is None:
        kernel = construct.wigner_kernel(L, N, reality, sampling, nside, forward)
    T = construct._ntheta(L, sampling, nside, forward)
    return _kernel_tensor(kernel, (N if reality else 2 * N - 1, T, L, 2 * L - 1), device)


def inverse(flmn, L, N, kernel=None, sampling="mw", reality=False, method="hip", nside=None):
    """Inverse Wigner transform via the precompute API (wigner.py:13-77)."""
    _check(method)
    s = sampling.lower()
    t, restore = sph._as_device_tensor(flmn, complex_only=True)
    with torch.cuda.device(t.device):
        K = _kernel(kernel, L, N, s, nside, reality, False, t.device)
        fshape = so3_samples.f_shape(L, N, s, nside)
        npix = int(np.prod(fshape[1:]))
        n0 = 0 if reality else -(N - 1)
        fban = torch.empty((N - n0,) + tuple(fshape[1:]), dtype=torch.complex128,
                           device=t.device)
        stream = sph._stream(t.device)
        for n in range(n0, N):
            plan = _plane_plan(s, L, nside, n, t.device)
            work = torch.empty(plan.workspace_bytes(1), dtype=torch.uint8, device=t.device)
            plane = t[N - 1 + n].contiguous()
            Kn = K[n - n0]
            _lib.check(_lib.lib().s2h_precompute_inverse(
                plan.handle, Kn.data_ptr(), int(Kn.is_complex()), 1.0, plane.data_ptr(),
                fban[n - n0].data_ptr(), work.data_ptr(), 1, stream), "precompute inverse")
        f = torch.empty(fshape, dtype=torch.float64 if reality else torch.complex128,
                        device=t.device)
        _lib.check(_lib.lib().s2h_gamma_dft(N, npix, 3 if reality else 1, 1.0, fban.data_ptr(),
                                            f.data_ptr(), stream), "gamma dft")
    return restore(f)


def forward(f, L, N, kernel=None, sampling="mw", reality=False, method="hip", nside=None):
    """Forward Wigner transform via the precompute API (wigner.py:224-288)."""
    _check(method)
    s = sampling.lower()
    t, restore = sph._as_device_tensor(f)
    with torch.cuda.device(t.device):
        K = _kernel(kernel, L, N, s, nside, reality, True, t.device)
        G = 2 * N - 1
        npix = t[0].numel()
        stream = sph._stream(t.device)
        if reality:
            src = (t.real if t.is_complex() else t).to(torch.float64).contiguous()
            mode = 2
        else:
            src = t.to(torch.complex128).contiguous()
            mode = 0
        # the kernel carries 2 pi / (2N-1): the gamma DFT is unnormalised here
        fban = torch.empty((G,) + tuple(t.shape[1:]), dtype=torch.complex128, device=t.device)
        _lib.check(_lib.lib().s2h_gamma_dft(N, npix, mode, 1.0, src.data_ptr(), fban.data_ptr(),
                                            stream), "gamma dft")
        flmn = torch.zeros((G, L, 2 * L - 1), dtype=torch.complex128, device=t.device)
        n0 = 0 if reality else -(N - 1)
        for n in range(n0, N):
            plan = _plane_plan(s, L, nside, n, t.device)
            work = torch.empty(plan.workspace_bytes(1), dtype=torch.uint8, device=t.device)
            Kn = K[n - n0]
            _lib.check(_lib.lib().s2h_precompute_forward(
                plan.handle, Kn.data_ptr(), int(Kn.is_complex()), 1.0,
                fban[N - 1 + n].data_ptr(), 1, flmn[N - 1 + n].data_ptr(), work.data_ptr(), 1,
                stream), "precompute forward")
        if reality:
            _lib.check(_lib.lib().s2h_wigner_reality_fill(flmn.data_ptr(), L, N, stream), "fill")
    return restore(flmn)
