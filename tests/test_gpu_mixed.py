"""Mixed-precision transforms (FP64 recursion, FP32 or split-BF16 matrix cores).

Tolerances are stated per test and precision.  "mixed": FP32 rounding of the
d-values and of the data plus FP32 accumulation over 128 rings (forward) or
32 l (inverse), about 1e-7 of max |flm| at these sizes; the tests allow 2e-6.
"bf16": hi + lo BF16 operands (16 significant bits), products without the
lo*lo term, FP32 accumulation, about 1e-5; the tests allow 1e-4.  "bf16x1"
(forward only): BF16 operands, one product, about 4e-3; the tests allow 2e-2.  Every test
also requires the error to be above 1e-12, which shows that the reduced
precision path (not the FP64 one) ran.  Refinement with an FP64 residual must
reach the FP64 target of 1e-7 against the seeded coefficients.
"""

import numpy as np
import pytest
import torch

from oracle import signals, transforms as otr
from tests.conftest import max_rel

from s2fft_amd.transforms import refine, spherical as sph

pytestmark = pytest.mark.gpu
MIXED_TOL = 2e-6
TOLS = {"mixed": 2e-6, "bf16": 1e-4, "bf16x1": 2e-2}


def _flms(rng, B, L, spin, reality):
    return np.stack([signals.generate_flm(rng, L, 0, spin, reality) for _ in range(B)])


CASES = [("healpix", 16, 32, 0, True, 3), ("gl", None, 24, 2, False, 2),
         ("mw", None, 16, 0, True, 5), ("mwss", None, 12, -1, False, 4),
         ("dh", None, 10, 0, False, 1)]


@pytest.mark.parametrize("precision", ["mixed", "bf16"])
@pytest.mark.parametrize("sampling,nside,L,spin,reality,B", CASES)
def test_mixed_inverse_vs_oracle(rng, sampling, nside, L, spin, reality, B, precision):
    flms = _flms(rng, B, L, spin, reality)
    got = sph.inverse_hip(torch.from_numpy(flms).cuda(), L, spin, nside, sampling, reality,
                          precision=precision).cpu().numpy()
    for i in range(B):
        ref = otr.inverse(flms[i], L, spin, nside, sampling, reality)
        if reality:
            ref = np.real(ref)
        err = max_rel(got[i], ref)
        assert 1e-12 < err < TOLS[precision], err


@pytest.mark.parametrize("precision", ["mixed", "bf16", "bf16x1"])
@pytest.mark.parametrize("sampling,nside,L,spin,reality,B", CASES)
def test_mixed_forward_vs_oracle(rng, sampling, nside, L, spin, reality, B, precision):
    flms = _flms(rng, B, L, spin, reality)
    fs = np.stack([otr.inverse(flms[i], L, spin, nside, sampling, reality) for i in range(B)])
    if reality:
        fs = np.real(fs)
    got = sph.forward_hip(torch.from_numpy(fs).cuda(), L, spin, nside, sampling, reality,
                          precision=precision).cpu().numpy()
    for i in range(B):
        ref = otr.forward(fs[i], L, spin, nside, sampling, reality)
        err = max_rel(got[i], ref)
        assert 1e-12 < err < TOLS[precision], err


@pytest.mark.parametrize("precision", ["mixed", "bf16"])
def test_mixed_vs_fp64_larger(rng, precision):
    """nside 128, L 256, 20 real maps (a partial map tile): against the FP64 kernels."""
    nside, L, B = 128, 256, 20
    flm = torch.from_numpy(_flms(rng, B, L, 0, True)).cuda()
    f64 = sph.inverse_hip(flm, L, 0, nside, "healpix", True)
    fmx = sph.inverse_hip(flm, L, 0, nside, "healpix", True, precision=precision)
    e_inv = float((fmx - f64).abs().max() / f64.abs().max())
    g64 = sph.forward_hip(f64, L, 0, nside, "healpix", True)
    gmx = sph.forward_hip(f64, L, 0, nside, "healpix", True, precision=precision)
    e_fwd = float((gmx - g64).abs().max() / g64.abs().max())
    print(f"{precision} vs fp64 at nside {nside}: inverse {e_inv:.2e}, forward {e_fwd:.2e}")
    assert 1e-12 < e_inv < TOLS[precision] and 1e-12 < e_fwd < TOLS[precision]


def test_bf16x1_forward_larger(rng):
    """nside 128, L 256, 20 real maps: single-product forward vs FP64."""
    nside, L, B = 128, 256, 20
    flm = torch.from_numpy(_flms(rng, B, L, 0, True)).cuda()
    f64 = sph.inverse_hip(flm, L, 0, nside, "healpix", True)
    g64 = sph.forward_hip(f64, L, 0, nside, "healpix", True)
    gx1 = sph.forward_hip(f64, L, 0, nside, "healpix", True, precision="bf16x1")
    e = float((gx1 - g64).abs().max() / g64.abs().max())
    print(f"bf16x1 forward vs fp64 at nside {nside}: {e:.2e}")
    assert 1e-12 < e < TOLS["bf16x1"]


def test_mixed_refinement_reaches_fp64_target(rng):
    """HEALPix nside 64, L 128, 4 real maps: refine_mixed reaches 1e-7."""
    nside, L, B = 64, 128, 4
    flm_true = torch.from_numpy(_flms(rng, B, L, 0, True)).cuda()
    f = sph.inverse_hip(flm_true, L, 0, nside, "healpix", True).real.contiguous()
    hist_m, hist_d = [], []
    refine.refine_mixed(f, L, 0, nside, "healpix", True, 10, history=hist_m, flm_true=flm_true)
    flm = sph.forward_hip(f, L, 0, nside, "healpix", True)
    for _ in range(10):
        flm = flm + sph.forward_hip(f - sph.inverse_hip(flm, L, 0, nside, "healpix", True).real,
                                    L, 0, nside, "healpix", True)
        hist_d.append(float((flm - flm_true).abs().max() / flm_true.abs().max()))
    print("mixed", ["%.1e" % e for e in hist_m])
    print("fp64 ", ["%.1e" % e for e in hist_d])
    assert min(hist_m) <= 1e-7
    # the mixed iteration contracts like the FP64 one (at most one step more)
    k = next(i for i, e in enumerate(hist_d) if e <= 1e-7)
    assert min(hist_m[k + 1:k + 3]) <= 1e-7


def test_bf16_refinement_with_refresh_reaches_fp64_target(rng):
    """Split-BF16 inner transforms with an FP64 residual every 3 steps reach 1e-7."""
    nside, L, B = 64, 128, 4
    flm_true = torch.from_numpy(_flms(rng, B, L, 0, True)).cuda()
    f = sph.inverse_hip(flm_true, L, 0, nside, "healpix", True).real.contiguous()
    hist = []
    refine.refine_mixed(f, L, 0, nside, "healpix", True, 12, history=hist, flm_true=flm_true,
                        precision="bf16", refresh_every=3)
    print("bf16+refresh", ["%.1e" % e for e in hist])
    assert min(hist) <= 1e-7


@pytest.mark.parametrize("shape,cplx", [((3, 1000), False), ((5, 7, 33), True), ((1, 300001), False),
                                        ((2, 0), False)])
def test_batch_sumsq(rng, shape, cplx):
    """s2h_batch_sumsq (the judge's residual norms) against numpy, and the
    same bits on a second call (fixed summation order)."""
    x = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0.0)
    xt = torch.from_numpy(x).cuda()
    got = refine._sumsq(xt).cpu().numpy()
    want = (np.abs(x.reshape(shape[0], -1)) ** 2).sum(axis=1)
    assert np.allclose(got, want, rtol=1e-13, atol=0.0)
    assert np.array_equal(refine._sumsq(xt).cpu().numpy(), got)


@pytest.mark.parametrize("reality", [True, False])
def test_bf16x1_forward_writes_every_entry(rng, reality):
    """The single-product BF16 forward writes zeros below l0 and the m < 0
    mirror itself (no zeroing or conjugate-fill pass): run it on memory the
    caching allocator hands back full of NaN, and check accumulate_into."""
    nside, L, B = 16, 32, 5
    flm_true = torch.from_numpy(_flms(rng, B, L, 0, reality)).cuda()
    f = sph.inverse_hip(flm_true, L, 0, nside, "healpix", reality)
    if reality:
        f = f.real.contiguous()
    junk = torch.full((B, L, 2 * L - 1), float("nan"), dtype=torch.complex128, device="cuda")
    del junk
    got = sph.forward_hip(f, L, 0, nside, "healpix", reality, precision="bf16x1")
    ref = sph.forward_hip(f, L, 0, nside, "healpix", reality)
    assert bool(torch.isfinite(got).all())
    e = max_rel(got.cpu().numpy(), ref.cpu().numpy())
    assert 1e-12 < e < TOLS["bf16x1"]
    # structural zeros exactly zero
    mask = np.abs(np.arange(-(L - 1), L))[None, :] > np.arange(L)[:, None]
    assert np.all(got.cpu().numpy()[:, mask] == 0)
    base = torch.from_numpy(_flms(rng, B, L, 0, reality)).cuda()
    acc = base.clone()
    sph.forward_hip(f, L, 0, nside, "healpix", reality, precision="bf16x1", accumulate_into=acc)
    assert max_rel(acc.cpu().numpy(), (base + got).cpu().numpy()) < 1e-15
