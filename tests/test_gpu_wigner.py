"""Wigner transforms: n-planes in one launch per stage, adjoints, autograd
(needs a GPU).

Follows the reference's tests/test_wigner_transform.py:37-79 (transform
parity, here against the oracle, which tests/test_oracle_pins.py pins by
brute-force D^l_mn synthesis) and tests/test_wigner_custom_grads.py:46,75
(check_grads, reverse mode): the hand-written adjoints are checked by
dot-product identities against the oracle's forward / inverse -- the
reference's gradients are the adjoints of exactly these maps -- and by
central finite differences (h = 1e-6), including the benchmark's
differentiation configuration (MW, L = 64, N = 4).
"""

import numpy as np
import pytest
import torch

from oracle import pm as opm
from oracle import signals, transforms as otr
from oracle import geometry as ogeo
from tests.conftest import max_rel

import s2fft_amd
from s2fft_amd import _lib
from s2fft_amd.transforms import wigner as wg

pytestmark = pytest.mark.gpu


class _nullctx:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


def _exact_if(sampling, N):
    return opm.exact_d() if (sampling == "healpix" and N > 1) else _nullctx()


def _rand(rng, shape, real):
    x = rng.standard_normal(shape)
    return x if real else x + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("reality", [False, True])
def test_planes_run_in_one_launch_per_stage(rng, reality):
    L, N = 6, 3
    flmn = signals.generate_flmn(rng, L, N, reality=reality)
    f = otr.wigner_inverse(flmn, L, N, sampling="mw", reality=reality)
    ft = torch.from_numpy(f).cuda()
    wg.forward_hip(ft, L, N, None, "mw", reality)          # plan creation outside the count
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    got = wg.forward_hip(ft, L, N, None, "mw", reality)
    fwd = {k: _lib.profile_query(k)[1] for k in (_lib.KERNEL_LAT_FORWARD,
                                                 _lib.KERNEL_RING_FFT_FORWARD,
                                                 _lib.KERNEL_RESAMPLE)}
    _lib.profile_enable(True)
    back = wg.inverse_hip(got, L, N, None, "mw", reality)
    inv = {k: _lib.profile_query(k)[1] for k in (_lib.KERNEL_LAT_INVERSE,
                                                 _lib.KERNEL_RING_FFT_INVERSE)}
    _lib.profile_enable(False)
    # complex: every stage once for all 2N-1 planes; reality: the n = 0 plane is
    # its own real transform (the reference's reality path), n >= 1 one batch
    want = 2 if reality else 1
    assert fwd == {_lib.KERNEL_LAT_FORWARD: want, _lib.KERNEL_RING_FFT_FORWARD: want,
                   _lib.KERNEL_RESAMPLE: want}
    assert inv == {_lib.KERNEL_LAT_INVERSE: want, _lib.KERNEL_RING_FFT_INVERSE: want}
    assert max_rel(got.cpu().numpy(), flmn) < 1e-11
    assert max_rel(back.cpu().numpy(), f) < 1e-12


CASES = [("mw", None, 6, 3, False), ("mw", None, 7, 2, True), ("mwss", None, 6, 2, False),
         ("mwss", None, 6, 3, True), ("gl", None, 5, 2, False), ("dh", None, 5, 3, True),
         ("healpix", 4, 8, 2, False), ("healpix", 4, 8, 3, True), ("mw", None, 5, 1, False),
         ("mw", None, 5, 1, True)]


@pytest.mark.parametrize("sampling,nside,L,N,reality", CASES)
def test_wigner_forward_adjoint_dot_product(rng, sampling, nside, L, N, reality):
    shape = ogeo.so3_f_shape(L, N, sampling, nside)
    x = _rand(rng, shape, reality)
    y = rng.standard_normal((2 * N - 1, L, 2 * L - 1)) + \
        1j * rng.standard_normal((2 * N - 1, L, 2 * L - 1))
    with _exact_if(sampling, N):
        Fx = otr.wigner_forward(x, L, N, nside, sampling, reality)
    lhs = np.real(np.vdot(y, Fx))
    FHy = wg.forward_adjoint_hip(torch.from_numpy(y).cuda(), L, N, nside, sampling,
                                 reality).cpu().numpy()
    rhs = np.real(np.vdot(FHy, x))
    assert abs(lhs - rhs) < 1e-11 * max(abs(lhs), 1.0), (lhs, rhs)


@pytest.mark.parametrize("sampling,nside,L,N,reality", CASES)
def test_wigner_inverse_adjoint_dot_product(rng, sampling, nside, L, N, reality):
    shape = ogeo.so3_f_shape(L, N, sampling, nside)
    flmn = rng.standard_normal((2 * N - 1, L, 2 * L - 1)) + \
        1j * rng.standard_normal((2 * N - 1, L, 2 * L - 1))
    for k in range(2 * N - 1):
        flmn[k, : abs(k - (N - 1))] = 0     # rows l < |n| are not coefficients
    g = _rand(rng, shape, reality)
    with _exact_if(sampling, N):
        If = otr.wigner_inverse(flmn, L, N, nside, sampling, reality)
    lhs = np.real(np.vdot(g, If))
    IHg = wg.inverse_adjoint_hip(torch.from_numpy(np.asarray(g)).cuda(), L, N, nside, sampling,
                                 reality).cpu().numpy()
    rhs = np.real(np.vdot(IHg, flmn))
    assert abs(lhs - rhs) < 1e-11 * max(abs(lhs), 1.0), (lhs, rhs)


def _fd_check(loss, x0, rng, real, n_dir=3):
    """Central differences (h = 1e-6) along random directions.  Bound: 1e-6
    relative plus the difference's rounding floor, 4 eps |loss| / h (the two
    losses are sums of O(|loss|) and lose that much in the subtraction)."""
    xt = torch.from_numpy(np.asarray(x0)).cuda().requires_grad_(True)
    l0 = loss(xt)
    l0.backward()
    grad = xt.grad.detach().cpu().numpy()
    h = 1e-6
    floor = 4 * np.finfo(np.float64).eps * abs(l0.item()) / h
    out = []
    for _ in range(n_dir):
        v = _rand(rng, np.shape(x0), real)
        vt = torch.from_numpy(np.asarray(v)).cuda()
        with torch.no_grad():
            fd = (loss(xt + h * vt) - loss(xt - h * vt)).item() / (2 * h)
        an = float(np.real(np.vdot(grad, v)))
        out.append((fd, an))
        assert abs(fd - an) < 1e-6 * abs(an) + floor, (fd, an, floor)
    return out


@pytest.mark.parametrize("reality", [False, True])
def test_wigner_gradients_finite_differences_small(rng, reality):
    L, N = 6, 3
    f0 = otr.wigner_inverse(signals.generate_flmn(rng, L, N, reality=reality), L, N,
                            reality=reality)
    t = torch.from_numpy(signals.generate_flmn(rng, L, N, reality=reality)).cuda()
    _fd_check(lambda ft: (s2fft_amd.wigner.forward(ft, L, N, reality=reality) - t).abs()
              .pow(2).sum(), np.real(f0) if reality else f0, rng, reality)
    ftarget = torch.from_numpy(np.asarray(f0)).cuda() * 0.5
    flmn0 = signals.generate_flmn(rng, L, N, reality=reality)
    _fd_check(lambda x: (s2fft_amd.wigner.inverse(x, L, N, reality=reality) - ftarget).abs()
              .pow(2).sum(), flmn0, rng, False)


def test_c4b_wigner_mw_L64_N4_gradients():
    """The differentiation configuration: MW, L = 64, N = 4, f (7, 64, 127) from
    seeded N(0,1) flmn; loss sum |forward(f) - t|^2 and sum |inverse(x) - g|^2,
    central finite differences (h = 1e-6) and the adjoint identities against
    the oracle's Wigner forward / inverse."""
    L, N = 64, 4
    rng = np.random.default_rng(400)
    flmn = signals.generate_flmn(np.random.default_rng(401), L, N)
    f0 = otr.wigner_inverse(flmn, L, N)
    assert f0.shape == (7, 64, 127)
    got_f = s2fft_amd.wigner.inverse(flmn, L, N, method="hip")
    e_inv = max_rel(got_f, f0)
    got = s2fft_amd.wigner.forward(f0, L, N, method="hip")
    e_fwd = max_rel(got, flmn)
    print(f"C4b: inverse vs oracle {e_inv:.2e}, forward round trip {e_fwd:.2e}", flush=True)
    assert e_inv < 1e-12 and e_fwd < 1e-11
    t = torch.from_numpy(signals.generate_flmn(np.random.default_rng(402), L, N)).cuda()
    fd = _fd_check(lambda ft: (s2fft_amd.wigner.forward(ft, L, N) - t).abs().pow(2).sum(),
                   f0, rng, False)
    print(f"C4b forward loss: FD vs adjoint {fd}", flush=True)
    g = torch.from_numpy(f0).cuda() * 0.5
    fd = _fd_check(lambda x: (s2fft_amd.wigner.inverse(x, L, N) - g).abs().pow(2).sum(),
                   flmn, rng, False)
    print(f"C4b inverse loss: FD vs adjoint {fd}", flush=True)
    # adjoint identities against the oracle at full size
    x = _rand(rng, f0.shape, False)
    y = _rand(rng, flmn.shape, False)
    lhs = np.real(np.vdot(y, otr.wigner_forward(x, L, N)))
    rhs = np.real(np.vdot(wg.forward_adjoint_hip(torch.from_numpy(y).cuda(), L, N)
                          .cpu().numpy(), x))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    lhs = np.real(np.vdot(x, otr.wigner_inverse(flmn, L, N)))
    rhs = np.real(np.vdot(wg.inverse_adjoint_hip(torch.from_numpy(x).cuda(), L, N)
                          .cpu().numpy(), flmn))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
