"""Adjoint kernels and autograd through the HIP transforms (needs a GPU).

Follows the reference's gradient tests (tests/test_spherical_custom_grads.py,
check_grads order 1, reverse mode): the hand-written adjoints are checked by
(i) dot-product identities against the CPU oracle's forward / inverse -- the
reference's custom VJPs plus JAX autodiff compute exactly these adjoints, so
this pins the gradients to the reference's definition -- and (ii) central
finite differences (h = 1e-6) of the loss sum |F(f) - t|^2.
"""

import numpy as np
import pytest
import torch

from oracle import pm as opm
from oracle import signals, transforms as otr
from oracle import geometry as ogeo

import s2fft_amd
from s2fft_amd.transforms import spherical as sph

pytestmark = pytest.mark.gpu

CASES = [("healpix", 8, 4, 0, True), ("healpix", 8, 4, 0, False), ("healpix", 9, 4, 1, False),
         ("gl", 7, None, 0, False), ("gl", 6, None, 0, True), ("dh", 6, None, 2, False),
         ("dh", 7, None, -1, False), ("mw", 6, None, 0, False), ("mw", 7, None, 0, True),
         ("mw", 9, None, -2, False), ("mwss", 6, None, 1, False), ("mwss", 7, None, 0, True),
         ("mw", 6, None, 0, True), ("mw", 7, None, 0, False)]


class _nullctx:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


def _exact_if(sampling, spin):
    """Spin != 0 HEALPix: exact d-values (the m-recursion drops NaN terms)."""
    return opm.exact_d() if (sampling == "healpix" and spin) else _nullctx()


def _rand_f(rng, L, sampling, nside, real):
    shape = ogeo.f_shape(L, sampling, nside)
    x = rng.standard_normal(shape)
    return x if real else x + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("sampling,L,nside,spin,reality", CASES)
def test_forward_adjoint_dot_product(rng, sampling, L, nside, spin, reality):
    x = _rand_f(rng, L, sampling, nside, reality)
    y = rng.standard_normal((L, 2 * L - 1)) + 1j * rng.standard_normal((L, 2 * L - 1))
    with _exact_if(sampling, spin):
        Fx = otr.forward(x, L, spin, nside, sampling, reality)
    lhs = np.real(np.vdot(y, Fx))
    yt = torch.from_numpy(y)[None].cuda()
    FHy = sph.forward_adjoint_hip(yt, L, spin, nside, sampling, reality)[0].cpu().numpy()
    rhs = np.real(np.vdot(FHy, x))
    assert abs(lhs - rhs) < 1e-11 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("sampling,L,nside,spin,reality", CASES)
def test_inverse_adjoint_dot_product(rng, sampling, L, nside, spin, reality):
    flm = rng.standard_normal((L, 2 * L - 1)) + 1j * rng.standard_normal((L, 2 * L - 1))
    flm[: abs(spin)] = 0      # rows l < |spin| are not harmonic coefficients
    g = _rand_f(rng, L, sampling, nside, reality)
    with _exact_if(sampling, spin):
        If = otr.inverse(flm, L, spin, nside, sampling, reality)
    lhs = np.real(np.vdot(g, If))
    gt = torch.from_numpy(np.asarray(g))[None].cuda()
    IHg = sph.inverse_adjoint_hip(gt, L, spin, nside, sampling, reality)[0].cpu().numpy()
    rhs = np.real(np.vdot(IHg, flm))
    assert abs(lhs - rhs) < 1e-11 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("sampling,L,nside,spin,reality", CASES[:4] + CASES[7:9])
def test_forward_loss_gradient_finite_differences(rng, sampling, L, nside, spin, reality):
    """Central FD, h = 1e-6, of sum |forward(f) - t|^2 (the Gauss-Legendre gradient case)."""
    f0 = _rand_f(rng, L, sampling, nside, reality)
    t = signals.generate_flm(rng, L, spin=spin, reality=reality)
    tt = torch.from_numpy(t).cuda()

    def loss(ft):
        return (s2fft_amd.forward(ft, L, spin, nside, sampling, "hip", reality) - tt).abs().pow(2).sum()

    ft = torch.from_numpy(np.asarray(f0)).cuda().requires_grad_(True)
    loss(ft).backward()
    grad = ft.grad.detach().cpu().numpy()
    h = 1e-6
    for _ in range(3):
        v = _rand_f(rng, L, sampling, nside, reality)
        vt = torch.from_numpy(np.asarray(v)).cuda()
        with torch.no_grad():
            fd = (loss(ft + h * vt) - loss(ft - h * vt)).item() / (2 * h)
        an = np.real(np.vdot(grad, v))
        assert abs(fd - an) < 1e-6 * max(abs(an), 1.0), (fd, an)


def test_refined_forward_gradient_flows(rng):
    """Gradients pass through the refinement loop (iter > 0)."""
    nside, L = 4, 8
    f0 = np.real(otr.inverse(signals.generate_flm(rng, L, reality=True), L, 0, nside,
                             "healpix", True))
    ft = torch.from_numpy(f0).cuda().requires_grad_(True)
    out = s2fft_amd.forward(ft, L, 0, nside, "healpix", "hip", True, iter=2)
    out.abs().pow(2).sum().backward()
    v = rng.standard_normal(f0.shape)
    vt = torch.from_numpy(v).cuda()
    h = 1e-6
    with torch.no_grad():
        lp = s2fft_amd.forward(ft + h * vt, L, 0, nside, "healpix", "hip", True, iter=2).abs().pow(2).sum()
        lm = s2fft_amd.forward(ft - h * vt, L, 0, nside, "healpix", "hip", True, iter=2).abs().pow(2).sum()
    fd = (lp - lm).item() / (2 * h)
    an = float((ft.grad * vt).sum())
    assert abs(fd - an) < 1e-6 * max(abs(an), 1.0)
