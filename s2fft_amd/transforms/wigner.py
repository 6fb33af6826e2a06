"""Wigner (SO(3)) transforms with ``method="hip"``.

Drop-in for s2fft.transforms.wigner.forward / inverse
(s2fft/transforms/wigner.py:13-107 and 346-439): f has shape [gamma, beta,
alpha] (or [gamma, pixel] for HEALPix) and flmn (2N-1, L, 2L-1).  A native
Wigner plan (``s2h_wigner_plan_create``) holds one spin -n recursion table
set per n; the gamma DFT, the ring FFTs, the theta resampling and the
latitudinal steps each run once over all n-planes (the reference maps a spin
transform over n with shared precomputes, price_mcewen.py:309-374, and
wigner.py:536-645).  Gradients: ``torch.autograd.Function`` wrappers whose
backward runs the hand-written adjoint transforms (``s2h_wigner_*_adjoint``),
recomputing the d-functions instead of storing them -- the reference
differentiates the same maps with JAX through its spin custom VJPs
(wigner.py:536-645, tests/test_wigner_custom_grads.py).
"""

import ctypes
import threading

import numpy as np
import torch

from .. import _lib
from ..sampling import so3_samples
from . import spherical as sph

METHODS = ("hip",)

_plan_lock = threading.Lock()
_plans = {}


class WignerDesc(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int32) for name in
                ("sampling", "L", "N", "nside", "L_lower", "reality")]


class WignerPlan:
    """Owner of one native Wigner plan (constant tables on the current device)."""

    def __init__(self, sampling, L, N, nside=None, L_lower=0, reality=False):
        s = sampling.lower()
        d = WignerDesc(_lib.SAMPLING_CODES[s], L, N, int(nside or 0) if s == "healpix" else 0,
                       L_lower, int(bool(reality)))
        handle = ctypes.c_void_p()
        _lib.check(_lib.lib().s2h_wigner_plan_create(ctypes.byref(handle), ctypes.byref(d)),
                   "wigner plan")
        self.handle = handle
        self.L, self.N, self.sampling, self.nside = L, N, s, nside
        self.reality = bool(reality)
        self.fshape = so3_samples.f_shape(L, N, s, nside)

    def workspace(self, device):
        n = int(_lib.lib().s2h_wigner_workspace_bytes(self.handle))
        return torch.empty(max(n, 8), dtype=torch.uint8, device=device)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.s2h_wigner_plan_destroy(h)
            self.handle = None


def get_plan(sampling, L, N, nside=None, L_lower=0, reality=False, device=None):
    """Cached native Wigner plan for the current (or given) device."""
    dev = torch.cuda.current_device() if device is None else torch.device(device).index
    if dev is None:
        dev = torch.cuda.current_device()
    s = sampling.lower()
    key = (dev, s, int(L), int(N), int(nside or 0) if s == "healpix" else 0, int(L_lower),
           bool(reality))
    with _plan_lock:
        plan = _plans.get(key)
        if plan is None:
            with torch.cuda.device(dev):
                plan = WignerPlan(s, L, N, nside, L_lower, reality)
            _plans[key] = plan
    return plan


def _check(method, N):
    if method not in METHODS:
        raise ValueError(f"Method {method} not recognised.")
    if N >= 8:
        raise Warning("Recursive transform may provide lower precision beyond N ~ 8")


def _check_shape(t, shape, what):
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what} has shape {tuple(t.shape)}, expected {tuple(shape)}")


# ------------------------------------------------------------------ device level
def forward_hip(f, L, N, nside=None, sampling="mw", reality=False, L_lower=0, plan=None):
    """One SO(3) signal on the device -> flmn (2N-1, L, 2L-1) complex128."""
    plan = plan or get_plan(sampling, L, N, nside, L_lower, reality, device=f.device)
    _check_shape(f, plan.fshape, "f")
    if reality:
        src = (f.real if f.is_complex() else f).to(torch.float64).contiguous()
    else:
        src = f.contiguous() if f.is_complex() else f.to(torch.float64).contiguous()
    flmn = torch.empty((2 * N - 1, L, 2 * L - 1), dtype=torch.complex128, device=f.device)
    work = plan.workspace(f.device)
    _lib.check(_lib.lib().s2h_wigner_forward(plan.handle, src.data_ptr(), int(src.is_complex()),
                                             flmn.data_ptr(), work.data_ptr(),
                                             sph._stream(f.device)), "wigner forward")
    return flmn


def inverse_hip(flmn, L, N, nside=None, sampling="mw", reality=False, L_lower=0, plan=None):
    """flmn (2N-1, L, 2L-1) on the device -> one SO(3) signal (float64 with
    reality, else complex128)."""
    plan = plan or get_plan(sampling, L, N, nside, L_lower, reality, device=flmn.device)
    _check_shape(flmn, (2 * N - 1, L, 2 * L - 1), "flmn")
    flmn = flmn.to(torch.complex128).contiguous()
    f = torch.empty(plan.fshape, dtype=torch.float64 if reality else torch.complex128,
                    device=flmn.device)
    work = plan.workspace(flmn.device)
    _lib.check(_lib.lib().s2h_wigner_inverse(plan.handle, flmn.data_ptr(), f.data_ptr(),
                                             work.data_ptr(), sph._stream(flmn.device)),
               "wigner inverse")
    return f


def forward_adjoint_hip(g, L, N, nside=None, sampling="mw", reality=False, L_lower=0, plan=None):
    """F^H g for a cotangent g shaped like flmn: float64 with reality, else complex128."""
    plan = plan or get_plan(sampling, L, N, nside, L_lower, reality, device=g.device)
    g = g.to(torch.complex128).contiguous()
    out = torch.empty(plan.fshape, dtype=torch.float64 if reality else torch.complex128,
                      device=g.device)
    work = plan.workspace(g.device)
    _lib.check(_lib.lib().s2h_wigner_forward_adjoint(plan.handle, g.data_ptr(), out.data_ptr(),
                                                     work.data_ptr(), sph._stream(g.device)),
               "wigner forward adjoint")
    return out


def inverse_adjoint_hip(g, L, N, nside=None, sampling="mw", reality=False, L_lower=0, plan=None):
    """I^H g for a cotangent g shaped like f (real with reality) -> (2N-1, L, 2L-1)."""
    plan = plan or get_plan(sampling, L, N, nside, L_lower, reality, device=g.device)
    if reality:
        g = (g.real if g.is_complex() else g).to(torch.float64).contiguous()
    else:
        g = g.contiguous() if g.is_complex() else g.to(torch.float64).contiguous()
    out = torch.empty((2 * N - 1, L, 2 * L - 1), dtype=torch.complex128, device=g.device)
    work = plan.workspace(g.device)
    _lib.check(_lib.lib().s2h_wigner_inverse_adjoint(plan.handle, g.data_ptr(),
                                                     int(g.is_complex()), out.data_ptr(),
                                                     work.data_ptr(), sph._stream(g.device)),
               "wigner inverse adjoint")
    return out


class _Forward(torch.autograd.Function):
    """flmn = F(f) with the hand-written adjoint as its backward (no stored d)."""

    @staticmethod
    def forward(ctx, f, L, N, nside, sampling, reality, L_lower):
        ctx.args = (L, N, nside, sampling, reality, L_lower)
        ctx.f_complex = f.is_complex()
        return forward_hip(f, L, N, nside, sampling, reality, L_lower)

    @staticmethod
    def backward(ctx, g):
        grad = forward_adjoint_hip(g, *ctx.args)
        if ctx.f_complex and not grad.is_complex():
            grad = grad.to(torch.complex128)
        elif not ctx.f_complex and grad.is_complex():
            grad = grad.real.contiguous()
        return grad, None, None, None, None, None, None


class _Inverse(torch.autograd.Function):
    """f = I(flmn) with the hand-written adjoint as its backward."""

    @staticmethod
    def forward(ctx, flmn, L, N, nside, sampling, reality, L_lower):
        ctx.args = (L, N, nside, sampling, reality, L_lower)
        return inverse_hip(flmn, L, N, nside, sampling, reality, L_lower)

    @staticmethod
    def backward(ctx, g):
        return (inverse_adjoint_hip(g, *ctx.args), None, None, None, None, None, None)


def forward_diff(f, L, N, nside=None, sampling="mw", reality=False, L_lower=0):
    if torch.is_grad_enabled() and f.requires_grad:
        return _Forward.apply(f, L, N, nside, sampling, reality, L_lower)
    return forward_hip(f, L, N, nside, sampling, reality, L_lower)


def inverse_diff(flmn, L, N, nside=None, sampling="mw", reality=False, L_lower=0):
    if torch.is_grad_enabled() and flmn.requires_grad:
        return _Inverse.apply(flmn, L, N, nside, sampling, reality, L_lower)
    return inverse_hip(flmn, L, N, nside, sampling, reality, L_lower)


# ------------------------------------------------------------------ reference API
def inverse(flmn, L, N, nside=None, sampling="mw", method="hip", reality=False,
            precomps=None, L_lower=0, _ssht_backend=1):
    """SO(3) inverse (s2fft/transforms/wigner.py:13-107)."""
    _check(method, N)
    t, restore = sph._as_device_tensor(flmn, complex_only=True)
    with torch.cuda.device(t.device):
        f = inverse_diff(t, L, N, nside, sampling, reality, L_lower)
    return restore(f)


def forward(f, L, N, nside=None, sampling="mw", method="hip", reality=False,
            precomps=None, L_lower=0, _ssht_backend=1):
    """SO(3) forward (s2fft/transforms/wigner.py:346-439)."""
    _check(method, N)
    t, restore = sph._as_device_tensor(f)
    with torch.cuda.device(t.device):
        flmn = forward_diff(t, L, N, nside, sampling, reality, L_lower)
    return restore(flmn)
