"""Precompute-path spin-spherical transforms with ``method="hip"``.

Drop-in for s2fft.precompute_transforms.spherical.inverse / forward
(spherical.py:19-84 and 226-311), with the same arguments, the reality
fallback warning for spin != 0 and the iterative refinement of the forward.
The transform is the reference's: the ring FFTs (+ MW/MWSS resampling)
without weights or phases, then the stored kernel applied as a dense
contraction over rings (forward, einsum "tlm,tm->lm", spherical.py:314-380)
or over l (inverse, "tlm,lm->tm", spherical.py:87-149), in HIP kernels
(``s2h_precompute_forward`` / ``s2h_precompute_inverse``).  ``kernel`` is the
tensor from :mod:`.construct` or a reference-made NumPy kernel of the same
layout; None builds it on the device.
"""

from warnings import warn

import numpy as np
import torch

from .. import _lib
from ..sampling import s2_samples
from ..transforms import spherical as sph
from . import construct

METHODS = ("hip",)


def _common(method, spin, reality):
    if method not in METHODS:
        raise ValueError(f"Method {method} not recognised.")
    if reality and spin != 0:
        warn("Reality acceleration only supports spin 0 fields. "
             "Defering to complex transform.", stacklevel=3)
        return False
    return reality


def _kernel_tensor(kernel, shape, device):
    """The kernel as a contiguous float64 / complex128 tensor on ``device``."""
    k = kernel if isinstance(kernel, torch.Tensor) else torch.from_numpy(np.asarray(kernel))
    if tuple(k.shape) != tuple(shape):
        raise ValueError(f"kernel has shape {tuple(k.shape)}, expected {tuple(shape)}")
    k = k.to(device)
    k = k.to(torch.complex128) if k.is_complex() else k.to(torch.float64)
    return k.contiguous()


def _kernel_shape(L, spin, sampling, nside, reality, forward):
    return (construct._ntheta(L, sampling, nside, forward), L, L if reality else 2 * L - 1)


def inverse_transform_hip(flm, kernel, L, sampling, reality, spin, nside):
    """flm (B, L, 2L-1) on the device -> f (B, ...) with a stored kernel."""
    s = sampling.lower()
    plan = sph.get_plan(s, L, nside, spin, 0, reality, device=flm.device)
    K = _kernel_tensor(kernel, _kernel_shape(L, spin, s, nside, reality, False), flm.device)
    B = flm.shape[0]
    cplx = plan.inverse_output_is_complex()
    f = torch.empty((B,) + s2_samples.f_shape(L, s, nside),
                    dtype=torch.complex128 if cplx else torch.float64, device=flm.device)
    work = torch.empty(plan.workspace_bytes(B), dtype=torch.uint8, device=flm.device)
    flm = flm.to(torch.complex128).contiguous()
    _lib.check(_lib.lib().s2h_precompute_inverse(plan.handle, K.data_ptr(), int(K.is_complex()),
                                                 float((-1) ** spin), flm.data_ptr(),
                                                 f.data_ptr(), work.data_ptr(), B,
                                                 sph._stream(flm.device)), "precompute inverse")
    return f


def forward_transform_hip(f, kernel, L, sampling, reality, spin, nside):
    """f (B, ...) on the device -> flm (B, L, 2L-1) with a stored kernel."""
    s = sampling.lower()
    plan = sph.get_plan(s, L, nside, spin, 0, reality, device=f.device)
    K = _kernel_tensor(kernel, _kernel_shape(L, spin, s, nside, reality, True), f.device)
    B = f.shape[0]
    flm = torch.empty((B, L, 2 * L - 1), dtype=torch.complex128, device=f.device)
    work = torch.empty(plan.workspace_bytes(B), dtype=torch.uint8, device=f.device)
    f = f.contiguous()
    _lib.check(_lib.lib().s2h_precompute_forward(plan.handle, K.data_ptr(), int(K.is_complex()),
                                                 float((-1) ** spin), f.data_ptr(),
                                                 int(f.is_complex()), flm.data_ptr(),
                                                 work.data_ptr(), B, sph._stream(f.device)),
               "precompute forward")
    return flm


def inverse(flm, L, spin=0, kernel=None, sampling="mw", reality=False, method="hip",
            nside=None):
    """Inverse transform via the precompute API (spherical.py:19-84)."""
    reality = _common(method, spin, reality)
    t, restore = sph._as_device_tensor(flm, complex_only=True)
    tb, batched = sph._split_batch(t, s2_samples.flm_shape(L), "flm")
    with torch.cuda.device(tb.device):
        if kernel is None:
            kernel = construct.spin_spherical_kernel(L, spin, reality, sampling, nside, False)
        f = inverse_transform_hip(tb, kernel, L, sampling, reality, spin, nside)
    return restore(f if batched else f[0])


def forward(f, L, spin=0, kernel=None, sampling="mw", reality=False, method="hip", nside=None,
            iter=0):
    """Forward transform via the precompute API (spherical.py:226-311); ``iter``
    steps of flm += F(f - I(flm)) with the stored kernels
    (s2fft/utils/iterative_refinement.py:9-46)."""
    reality = _common(method, spin, reality)
    t, restore = sph._as_device_tensor(f)
    tb, batched = sph._split_batch(t, s2_samples.f_shape(L, sampling, nside), "f")
    with torch.cuda.device(tb.device):
        if kernel is None:
            kernel = construct.spin_spherical_kernel(L, spin, reality, sampling, nside, True)
        flm = forward_transform_hip(tb, kernel, L, sampling, reality, spin, nside)
        if iter:
            inv_kernel = construct.spin_spherical_kernel(L, spin, reality, sampling, nside, False)
            for _ in range(iter):
                rec = inverse_transform_hip(flm, inv_kernel, L, sampling, reality, spin, nside)
                if not tb.is_complex() and rec.is_complex():
                    rec = rec.real
                flm = flm + forward_transform_hip(tb - rec, kernel, L, sampling, reality, spin,
                                                  nside)
    return restore(flm if batched else flm[0])
