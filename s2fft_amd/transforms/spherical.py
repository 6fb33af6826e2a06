"""Spin spherical-harmonic transforms with ``method="hip"``.

Drop-in for s2fft.transforms.spherical.forward / inverse
(s2fft/transforms/spherical.py:21-109 and 352-464): same positional and
keyword arguments, same array layouts (flm is (L, 2L-1) with column L-1+m,
f as s2fft.sampling.s2_samples.f_shape), same exceptions.  Every transform
runs on the GPU through the C ABI of ``libs2fft_amd.so``; there is no CPU
path.  Inputs may be NumPy arrays (the result is returned as NumPy) or torch
tensors (the result stays on the tensor's device); a leading batch axis is
accepted and transformed in one launch per kernel.
"""

import ctypes
import threading

import numpy as np
import torch

from .. import _lib
from ..sampling import s2_samples as samples

METHODS = ("hip",)

_plan_lock = threading.Lock()
_plans = {}


def get_plan(sampling, L, nside=None, spin=0, L_lower=0, reality=False,
             norm=_lib.NORM_SPIN, extra_sign=1, device=None):
    """Cached native plan for the current (or given) CUDA/HIP device."""
    dev = torch.cuda.current_device() if device is None else torch.device(device).index
    if dev is None:
        dev = torch.cuda.current_device()
    s = sampling.lower()
    key = (dev, s, int(L), int(nside or 0) if s == "healpix" else 0, int(spin), int(L_lower),
           bool(reality), int(norm), int(extra_sign))
    with _plan_lock:
        plan = _plans.get(key)
        if plan is None:
            with torch.cuda.device(dev):
                plan = _lib.Plan(s, L, nside if s == "healpix" else 0, spin, L_lower, reality,
                                 norm, extra_sign)
            _plans[key] = plan
    return plan


def clear_plans():
    with _plan_lock:
        _plans.clear()


# ------------------------------------------------------------------ helpers
def _as_device_tensor(x, complex_only=False):
    """(tensor on a GPU, restore) where restore maps a result back to x's kind."""
    if isinstance(x, torch.Tensor):
        if x.is_cuda:
            t, restore = x, (lambda y: y)
        else:
            t = x.to("cuda")
            restore = (lambda y: y.cpu())
    else:
        t = torch.from_numpy(np.ascontiguousarray(x)).to("cuda")
        restore = (lambda y: y.cpu().numpy())
    if t.is_complex():
        t = t.to(torch.complex128)
    elif complex_only:
        t = t.to(torch.complex128)
    else:
        t = t.to(torch.float64)
    return t.contiguous(), restore


def _stream(device):
    return torch.cuda.current_stream(device).cuda_stream


def _split_batch(t, core_shape, what):
    core = tuple(core_shape)
    if tuple(t.shape) == core:
        return t.reshape((1,) + core), False
    if t.dim() == len(core) + 1 and tuple(t.shape[1:]) == core:
        return t, True
    raise ValueError(f"{what} has shape {tuple(t.shape)}, expected {core} or (batch,) + {core}")


def _check_common(method, spin, sampling, nside):
    if method not in METHODS:
        raise ValueError(f"Method {method} not recognised.")
    if spin >= 8:
        raise Warning("Recursive transform may provide lower precision beyond spin ~ 8")
    if sampling.lower() not in samples.SAMPLINGS:
        raise ValueError(f"Sampling scheme sampling={sampling} not supported")
    if sampling.lower() == "healpix" and nside is None:
        raise ValueError(f"Sampling scheme sampling={sampling} with nside={nside} not supported")


PRECISIONS = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED, "bf16": _lib.PREC_BF16,
              "bf16x1": _lib.PREC_BF16X1}


def _precision(name):
    try:
        return PRECISIONS[name]
    except KeyError:
        raise ValueError(f"precision={name!r}: expected one of {sorted(PRECISIONS)}") from None


# ------------------------------------------------------------------ device level
def inverse_hip(flm, L, spin=0, nside=None, sampling="mw", reality=False, L_lower=0,
                plan=None, precision="fp64", real_output=False, sub_from=None):
    """Batched inverse on device tensors: flm (B, L, 2L-1) complex128 -> f (B, ...).

    ``precision``: "fp64" (default), "mixed" (FP64 recursion, FP32 matrix-core
    synthesis with FP32 accumulation within 32-l chunks) or "bf16" (FP64
    recursion, split-BF16 matrix-core synthesis, FP32 accumulation).
    ``real_output`` (reality only): float64 pixels even where the reference
    returns complex128 with a zero imaginary part (HEALPix).  ``sub_from``:
    return sub_from - I(flm) (a residual) in the same pass; it may be
    overwritten in place when it is the returned tensor's only owner.
    """
    s = sampling.lower()
    plan = plan or get_plan(s, L, nside, spin, L_lower, reality, device=flm.device)
    B = flm.shape[0]
    fshape = samples.f_shape(L, s, nside)
    cplx_out = plan.inverse_output_is_complex() and not real_output
    dtype = torch.complex128 if cplx_out else torch.float64
    flags = _lib.INV_REAL_OUTPUT if real_output else 0
    if sub_from is not None:
        sub_from = sub_from.to(dtype).contiguous()
        if tuple(sub_from.shape) != (B,) + tuple(fshape):
            raise ValueError(f"sub_from has shape {tuple(sub_from.shape)}")
        flags |= _lib.INV_SUBTRACT
    f = torch.empty((B,) + fshape, dtype=dtype, device=flm.device)
    work = torch.empty(plan.workspace_bytes(B), dtype=torch.uint8, device=flm.device)
    flm = flm.contiguous()
    _lib.check(_lib.lib().s2h_inverse_ex(
        plan.handle, flm.data_ptr(), f.data_ptr(), work.data_ptr(), B, _precision(precision),
        flags, sub_from.data_ptr() if sub_from is not None else None, _stream(flm.device)),
        "inverse")
    return f


def forward_hip(f, L, spin=0, nside=None, sampling="mw", reality=False, L_lower=0, plan=None,
                precision="fp64", accumulate_into=None):
    """Batched forward on device tensors: f (B, ...) -> flm (B, L, 2L-1) complex128.

    ``precision``: "fp64" (default), "mixed" (FP64 recursion and ring FFTs,
    FP32 matrix-core projection with FP32 accumulation over 128 rings),
    "bf16" (split-BF16 matrix-core projection, FP32 accumulation) or "bf16x1"
    (single-product BF16 projection, ~4e-3 relative: a refinement's
    correction transform).  ``accumulate_into``: an flm tensor (contiguous
    complex128, same shape) that receives flm += F(f) in place and is returned.
    """
    s = sampling.lower()
    plan = plan or get_plan(s, L, nside, spin, L_lower, reality, device=f.device)
    B = f.shape[0]
    flags = 0
    if accumulate_into is not None:
        flm = accumulate_into
        if (flm.dtype != torch.complex128 or not flm.is_contiguous()
                or tuple(flm.shape) != (B, L, 2 * L - 1)):
            raise ValueError("accumulate_into must be a contiguous (B, L, 2L-1) complex128 tensor")
        flags = _lib.FWD_ACCUMULATE
    else:
        flm = torch.empty((B, L, 2 * L - 1), dtype=torch.complex128, device=f.device)
    work = torch.empty(plan.workspace_bytes(B), dtype=torch.uint8, device=f.device)
    f = f.contiguous()
    _lib.check(_lib.lib().s2h_forward_ex(plan.handle, f.data_ptr(), int(f.is_complex()),
                                         flm.data_ptr(), work.data_ptr(), B,
                                         _precision(precision), flags, _stream(f.device)),
               "forward")
    return flm


def forward_adjoint_hip(g, L, spin=0, nside=None, sampling="mw", reality=False, L_lower=0,
                        plan=None):
    """F^H g on device tensors (g: (B, L, 2L-1) complex128)."""
    s = sampling.lower()
    plan = plan or get_plan(s, L, nside, spin, L_lower, reality, device=g.device)
    B = g.shape[0]
    dtype = torch.float64 if reality else torch.complex128
    out = torch.empty((B,) + samples.f_shape(L, s, nside), dtype=dtype, device=g.device)
    work = torch.empty(plan.workspace_bytes(B), dtype=torch.uint8, device=g.device)
    g = g.contiguous().to(torch.complex128)
    _lib.check(_lib.lib().s2h_forward_adjoint(plan.handle, g.data_ptr(), out.data_ptr(),
                                              work.data_ptr(), B, _stream(g.device)),
               "forward adjoint")
    return out


def inverse_adjoint_hip(g, L, spin=0, nside=None, sampling="mw", reality=False, L_lower=0,
                        plan=None):
    """I^H g on device tensors (g in the f layout) -> (B, L, 2L-1) complex128."""
    s = sampling.lower()
    plan = plan or get_plan(s, L, nside, spin, L_lower, reality, device=g.device)
    B = g.shape[0]
    out = torch.empty((B, L, 2 * L - 1), dtype=torch.complex128, device=g.device)
    work = torch.empty(plan.workspace_bytes(B), dtype=torch.uint8, device=g.device)
    g = g.contiguous()
    if not g.is_complex():
        g = g.to(torch.float64)
    _lib.check(_lib.lib().s2h_inverse_adjoint(plan.handle, g.data_ptr(), int(g.is_complex()),
                                              out.data_ptr(), work.data_ptr(), B,
                                              _stream(g.device)), "inverse adjoint")
    return out


class _Forward(torch.autograd.Function):
    """flm = F(f) with the hand-written adjoint as its backward (no stored d)."""

    @staticmethod
    def forward(ctx, f, L, spin, nside, sampling, reality, L_lower):
        ctx.args = (L, spin, nside, sampling, reality, L_lower)
        ctx.f_complex = f.is_complex()
        return forward_hip(f, L, spin, nside, sampling, reality, L_lower)

    @staticmethod
    def backward(ctx, g):
        grad = forward_adjoint_hip(g, *ctx.args)
        if ctx.f_complex and not grad.is_complex():
            grad = grad.to(torch.complex128)
        elif not ctx.f_complex and grad.is_complex():
            grad = grad.real.contiguous()
        return grad, None, None, None, None, None, None


class _Inverse(torch.autograd.Function):
    """f = I(flm) with the hand-written adjoint as its backward."""

    @staticmethod
    def forward(ctx, flm, L, spin, nside, sampling, reality, L_lower):
        ctx.args = (L, spin, nside, sampling, reality, L_lower)
        return inverse_hip(flm, L, spin, nside, sampling, reality, L_lower)

    @staticmethod
    def backward(ctx, g):
        return (inverse_adjoint_hip(g, *ctx.args), None, None, None, None, None, None)


def forward_diff(f, L, spin=0, nside=None, sampling="mw", reality=False, L_lower=0):
    """forward_hip that records the adjoint for autograd when needed."""
    if torch.is_grad_enabled() and f.requires_grad:
        return _Forward.apply(f, L, spin, nside, sampling, reality, L_lower)
    return forward_hip(f, L, spin, nside, sampling, reality, L_lower)


def inverse_diff(flm, L, spin=0, nside=None, sampling="mw", reality=False, L_lower=0):
    if torch.is_grad_enabled() and flm.requires_grad:
        return _Inverse.apply(flm, L, spin, nside, sampling, reality, L_lower)
    return inverse_hip(flm, L, spin, nside, sampling, reality, L_lower)


# ------------------------------------------------------------------ ring spectra
# The transforms split at the ring spectra (s2h_ring_* in the C ABI): with
# G = ring_analysis_hip(f), forward = forward_rings_hip(G) and
# F(f - I(x)) = forward_rings_hip(ring_residual_hip(G, inverse_rings_hip(x))),
# where the residual applies (ring FFT . ring iFFT) without either FFT.
# ``tiled``: spectra in blocks of 8 rows (S2H_SPECTRA_TILED), the layout in
# which the latitudinal kernels' column gathers and scatters touch whole
# cache lines; ring_analysis_hip writes ring-major spectra.
def spectra_supported(plan):
    """Whether the plan's forward and inverse share one ring grid (HEALPix, GL)."""
    return int(_lib.lib().s2h_ring_spectra_bytes(plan.handle, 1)) > 0


def _new_spectra(plan, B, device):
    nbytes = int(_lib.lib().s2h_ring_spectra_bytes(plan.handle, B))
    if nbytes == 0 and B > 0:
        raise ValueError("ring spectra need a plan whose forward and inverse share one ring grid")
    return torch.empty((B, nbytes // (16 * max(B, 1))), dtype=torch.complex128, device=device)


def ring_analysis_hip(f, plan):
    """Ring spectra G (B, rows * Mc) complex128 of the maps f (B, ...)."""
    f = f.contiguous()
    G = _new_spectra(plan, f.shape[0], f.device)
    _lib.check(_lib.lib().s2h_ring_analysis(plan.handle, f.data_ptr(), int(f.is_complex()),
                                            G.data_ptr(), f.shape[0], _stream(f.device)),
               "ring analysis")
    return G


def _flm_shape(B, L, packed):
    return (B, L * (L + 1) // 2) if packed else (B, L, 2 * L - 1)


def forward_rings_hip(G, L, plan, precision="fp64", accumulate_into=None, packed=False,
                      tiled=False):
    """flm (B, L, 2L-1) from ring spectra; ``accumulate_into``: flm += in place.
    ``packed``: flm in the packed m-major layout (B, L(L+1)/2) (real maps).
    ``tiled``: G in the tiled layout."""
    B = G.shape[0]
    flags = (_lib.FLM_PACKED if packed else 0) | (_lib.SPECTRA_TILED if tiled else 0)
    shape = _flm_shape(B, L, packed)
    if accumulate_into is not None:
        flm = accumulate_into
        if (flm.dtype != torch.complex128 or not flm.is_contiguous()
                or tuple(flm.shape) != shape):
            raise ValueError(f"accumulate_into must be a contiguous {shape} complex128 tensor")
        flags |= _lib.FWD_ACCUMULATE
    else:
        flm = torch.empty(shape, dtype=torch.complex128, device=G.device)
    _lib.check(_lib.lib().s2h_forward_rings(plan.handle, G.contiguous().data_ptr(),
                                            flm.data_ptr(), B, _precision(precision), flags,
                                            _stream(G.device)), "forward rings")
    return flm


def inverse_rings_hip(flm, plan, precision="fp64", packed=False, tiled=False):
    """Ring spectra (B, rows * Mc) of the latitudinal synthesis of flm
    (``packed``: flm in the packed m-major layout; ``tiled``: the spectra
    in the tiled layout)."""
    flm = flm.contiguous()
    Y = _new_spectra(plan, flm.shape[0], flm.device)
    flags = (_lib.FLM_PACKED if packed else 0) | (_lib.SPECTRA_TILED if tiled else 0)
    _lib.check(_lib.lib().s2h_inverse_rings(plan.handle, flm.data_ptr(), Y.data_ptr(),
                                            flm.shape[0], _precision(precision), flags,
                                            _stream(flm.device)), "inverse rings")
    return Y


def flm_pack_hip(flm, L):
    """(B, L, 2L-1) -> packed m-major (B, L(L+1)/2) (real maps: m >= 0 kept)."""
    flm = flm.contiguous()
    out = torch.empty(_flm_shape(flm.shape[0], L, True), dtype=torch.complex128,
                      device=flm.device)
    _lib.check(_lib.lib().s2h_flm_pack(L, flm.data_ptr(), out.data_ptr(), flm.shape[0],
                                       _stream(flm.device)), "flm pack")
    return out


def flm_unpack_hip(flm_p, L):
    """Packed m-major (B, L(L+1)/2) -> (B, L, 2L-1) with the m < 0 half."""
    flm_p = flm_p.contiguous()
    out = torch.empty(_flm_shape(flm_p.shape[0], L, False), dtype=torch.complex128,
                      device=flm_p.device)
    _lib.check(_lib.lib().s2h_flm_unpack(L, flm_p.data_ptr(), out.data_ptr(), flm_p.shape[0],
                                         _stream(flm_p.device)), "flm unpack")
    return out


def ring_residual_hip(G, Y, plan, out=None, alpha=None, sumsq=False, tiled=False,
                      g_tiled=None):
    """R = G - A(Y), or G - alpha[b] A(Y) with a per-map float64 device tensor
    ``alpha`` (out may be G).  ``sumsq``: also return the per-map sums of
    |R|^2 (one pass, fixed order) as (R, sumsq).  ``tiled``: Y and R in the
    tiled layout (real maps), G too unless ``g_tiled`` is False (G as
    ring_analysis_hip writes it; out may then not be G)."""
    B = G.shape[0]
    R = torch.empty_like(G) if out is None else out
    a = None
    if alpha is not None:
        if alpha.dtype != torch.float64 or alpha.numel() != B or alpha.device != G.device:
            raise ValueError("alpha: one float64 per map on the spectra's device")
        a = alpha.contiguous().data_ptr()
    sq = work = None
    lib = _lib.lib()
    if sumsq:
        sq = torch.empty(B, dtype=torch.float64, device=G.device)
        work = torch.empty(max(int(lib.s2h_ring_dot_workspace_bytes(plan.handle, B)), 8),
                           dtype=torch.uint8, device=G.device)
    flags = 0
    if tiled:
        flags = _lib.SPECTRA_TILED | (0 if g_tiled in (None, True) else _lib.SPECTRA_G_RING_MAJOR)
    _lib.check(lib.s2h_ring_residual_ex(plan.handle, G.contiguous().data_ptr(),
                                        Y.contiguous().data_ptr(), a, R.data_ptr(),
                                        sq.data_ptr() if sq is not None else None,
                                        work.data_ptr() if work is not None else None,
                                        B, flags, _stream(G.device)), "ring residual")
    return (R, sq) if sumsq else R


def ring_dot_hip(Y, plan, num=None, tiled=False):
    """pq[b] = <p, F A(Y)> for Y = inverse_rings_hip(p) (the curvature of a
    conjugate-gradient step); with ``num`` also alpha = num / pq.  ``tiled``:
    Y in the tiled layout (real maps).  Returns pq, or (pq, alpha)."""
    B = Y.shape[0]
    lib = _lib.lib()
    pq = torch.empty(B, dtype=torch.float64, device=Y.device)
    alpha = torch.empty_like(pq) if num is not None else None
    work = torch.empty(max(int(lib.s2h_ring_dot_workspace_bytes(plan.handle, B)), 8),
                       dtype=torch.uint8, device=Y.device)
    _lib.check(lib.s2h_ring_dot_ex(plan.handle, Y.contiguous().data_ptr(),
                                   num.contiguous().data_ptr() if num is not None else None,
                                   pq.data_ptr(), alpha.data_ptr() if alpha is not None else None,
                                   work.data_ptr(), B, _lib.SPECTRA_TILED if tiled else 0,
                                   _stream(Y.device)), "ring dot")
    return pq if num is None else (pq, alpha)


def inverse_rings_dot_hip(flm, plan, precision="fp64", packed=False, num=None):
    """Y = inverse_rings_hip(flm, tiled=True) and ring_dot_hip(Y, num=num,
    tiled=True) in one call (real maps): the matrix-core syntheses add up the
    dot over the rows without aliasing as they store them.  Returns (Y, pq)
    or, with ``num``, (Y, pq, alpha)."""
    flm = flm.contiguous()
    B = flm.shape[0]
    lib = _lib.lib()
    Y = _new_spectra(plan, B, flm.device)
    pq = torch.empty(B, dtype=torch.float64, device=flm.device)
    alpha = torch.empty_like(pq) if num is not None else None
    work = torch.empty(max(int(lib.s2h_inverse_rings_dot_workspace_bytes(plan.handle, B)), 8),
                       dtype=torch.uint8, device=flm.device)
    flags = _lib.SPECTRA_TILED | (_lib.FLM_PACKED if packed else 0)
    _lib.check(lib.s2h_inverse_rings_dot(plan.handle, flm.data_ptr(), Y.data_ptr(), B,
                                         _precision(precision), flags,
                                         num.contiguous().data_ptr() if num is not None else None,
                                         pq.data_ptr(),
                                         alpha.data_ptr() if alpha is not None else None,
                                         work.data_ptr(), _stream(flm.device)),
               "inverse rings + dot")
    return (Y, pq) if num is None else (Y, pq, alpha)


def flm_dots_hip(z, vs, packed_L=0):
    """(len(vs), B) per-map inner products sum w Re(conj(z) v) of coefficient
    tensors (B, ...) complex128; ``packed_L`` > 0: the packed layout of real
    maps (weight 2 beyond the m = 0 block)."""
    z = z.contiguous()
    B = z.shape[0]
    n = z[0].numel() if B else 0
    vs = [v.contiguous() for v in vs]
    for v in vs:
        if v.shape != z.shape or v.dtype != torch.complex128:
            raise ValueError("flm_dots: vectors must match z (complex128)")
    lib = _lib.lib()
    out = torch.empty((len(vs), B), dtype=torch.float64, device=z.device)
    work = torch.empty(max(int(lib.s2h_flm_dots_workspace_bytes(len(vs), B)), 8),
                       dtype=torch.uint8, device=z.device)
    ptrs = (ctypes.c_void_p * len(vs))(*[v.data_ptr() for v in vs])
    _lib.check(lib.s2h_flm_dots(z.data_ptr(), ptrs, len(vs), n, int(packed_L), B,
                                out.data_ptr(), work.data_ptr(), _stream(z.device)), "flm dots")
    return out


def cg_update_hip(x, p, alpha, p_new=None, z=None, g1=None, p_prev=None, g2=None):
    """x += alpha p (x may be None), and p_new = z - g1 p - g2 p_prev (per-map
    float64 scalars on the device; p_prev optional; p_new may be z)."""
    B = p.shape[0]
    n = p[0].numel() if B else 0
    for t in (x, p, p_new, z, p_prev):
        if t is not None and (not t.is_contiguous() or t.shape != p.shape):
            raise ValueError("cg_update: contiguous tensors of one shape")

    def ptr(t):
        return t.data_ptr() if t is not None else None

    _lib.check(_lib.lib().s2h_cg_update(ptr(x), ptr(p), ptr(alpha), ptr(p_new), ptr(z), ptr(g1),
                                        ptr(p_prev), ptr(g2), n, B, _stream(p.device)),
               "cg update")
    return x


def cg_finish_hip(x, ps, alphas, unpack_L=None):
    """x + sum_j alphas[j] ps[j] (per-map float64 scalars, at most 4 terms, added
    in order): in place into x, or, with ``unpack_L``, from the packed layout of
    real maps into a new (B, L, 2L-1) tensor in the reference layout."""
    B = x.shape[0]
    n = x[0].numel() if B else 0
    if len(ps) != len(alphas) or len(ps) > 4:
        raise ValueError("cg_finish: up to 4 (direction, step) pairs")
    for t in ps:
        if not t.is_contiguous() or t.shape != x.shape:
            raise ValueError("cg_finish: contiguous directions shaped like x")
    out = None
    if unpack_L is not None:
        out = torch.empty(_flm_shape(B, unpack_L, False), dtype=torch.complex128, device=x.device)
    pp = (ctypes.c_void_p * max(len(ps), 1))(*[t.data_ptr() for t in ps])
    aa = (ctypes.c_void_p * max(len(ps), 1))(*[a.contiguous().data_ptr() for a in alphas])
    _lib.check(_lib.lib().s2h_cg_finish(x.data_ptr(), pp, aa, len(ps), n,
                                        int(unpack_L or 0), out.data_ptr() if out is not None
                                        else None, B, _stream(x.device)), "cg finish")
    return out if out is not None else x


def forward_iterative_hip(f, L, spin=0, nside=None, sampling="mw", reality=False, L_lower=0,
                          n_iter=0, spectral=True):
    """flm = F(f); repeat n_iter: flm += F(f - I(flm)).

    s2fft/utils/iterative_refinement.py:9-46 as used by spherical.py:453-462;
    the residual and the update stay in FP64 on the device.  Without autograd,
    on a plan with one ring grid (HEALPix) and ``spectral``, the residual is
    formed on the ring spectra, F(f - I(flm)) = Lat_fwd(G - A(Lat_inv(flm))),
    so each iteration runs no ring FFT; otherwise the residual comes out of the
    synthesis pass (f - I(flm)) and the update is accumulated by the
    projection pass (flm += F(r)).
    """
    if n_iter and not (torch.is_grad_enabled() and f.requires_grad):
        s = sampling.lower()
        plan = get_plan(s, L, nside, spin, L_lower, reality, device=f.device)
        f = f.contiguous()
        if not f.is_complex():
            f = f.to(torch.float64)
        if spectral and spectra_supported(plan):
            # real maps iterate on the packed m-major coefficients
            pk = bool(reality)
            G = ring_analysis_hip(f, plan)
            flm = forward_rings_hip(G, L, plan, packed=pk)
            for _ in range(n_iter):
                R = ring_residual_hip(G, inverse_rings_hip(flm, plan, packed=pk), plan)
                forward_rings_hip(R, L, plan, accumulate_into=flm, packed=pk)
            return flm_unpack_hip(flm, L) if pk else flm
        flm = forward_hip(f, L, spin, nside, s, reality, L_lower, plan=plan)
        real_out = bool(reality) and not f.is_complex()
        for _ in range(n_iter):
            r = inverse_hip(flm, L, spin, nside, s, reality, L_lower, plan=plan,
                            real_output=real_out, sub_from=f)
            forward_hip(r, L, spin, nside, s, reality, L_lower, plan=plan,
                        accumulate_into=flm)
        return flm
    flm = forward_diff(f, L, spin, nside, sampling, reality, L_lower)
    for _ in range(n_iter):
        f_rec = inverse_diff(flm, L, spin, nside, sampling, reality, L_lower)
        flm = flm + forward_diff(f - f_rec, L, spin, nside, sampling, reality, L_lower)
    return flm


# ------------------------------------------------------------------ reference API
def inverse(flm, L, spin=0, nside=None, sampling="mw", method="hip", reality=False,
            precomps=None, spmd=False, L_lower=0, _ssht_backend=1):
    """Inverse spin-spherical harmonic transform (s2fft/transforms/spherical.py:21-109).

    ``precomps``, ``spmd`` and ``_ssht_backend`` are accepted for signature
    compatibility; the recursion tables live in a cached native plan and the
    transform runs on the current GPU.
    """
    _check_common(method, spin, sampling, nside)
    t, restore = _as_device_tensor(flm, complex_only=True)
    tb, batched = _split_batch(t, samples.flm_shape(L), "flm")
    with torch.cuda.device(tb.device):
        f = inverse_diff(tb, L, spin, nside, sampling, reality, L_lower)
    return restore(f if batched else f[0])


def forward(f, L, spin=0, nside=None, sampling="mw", method="hip", reality=False,
            precomps=None, spmd=False, L_lower=0, iter=None, _ssht_backend=1):
    """Forward spin-spherical harmonic transform (s2fft/transforms/spherical.py:352-464).

    ``iter`` > 0 applies that many steps of iterative refinement
    (default 0, as for every method except jax_healpy in the reference).
    """
    _check_common(method, spin, sampling, nside)
    if iter is None:
        iter = 0
    t, restore = _as_device_tensor(f)
    tb, batched = _split_batch(t, samples.f_shape(L, sampling, nside), "f")
    with torch.cuda.device(tb.device):
        flm = forward_iterative_hip(tb, L, spin, nside, sampling, reality, L_lower, iter)
    return restore(flm if batched else flm[0])
