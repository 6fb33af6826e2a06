"""Ring-spectra split of the transforms (s2h_ring_* in the C ABI).

forward = forward_rings . ring_analysis must hold bit for bit (the same
kernels run in the same order).  The spectral residual G - A(Lat_inv(x))
equals the ring spectra of the pixel residual f - I(x) up to rounding: the
two paths differ only by a pair of ring FFTs, so the tolerance is 1e-12 of
max |G|.  Refinement in spectral space must follow the pixel iteration to the
same tolerance (FP64) and reach the 1e-7 target (mixed precision).  The
HEALPix sizes include Bluestein rings (n_t = 4 i with a prime factor of i
above 13, from nside 18 on) and rings shorter than 2L (folded spectra).
"""

import numpy as np
import pytest
import torch

from oracle import signals
from s2fft_amd.transforms import refine, spherical as sph

pytestmark = pytest.mark.gpu

CASES = [("healpix", 8, 16, True), ("healpix", 8, 24, False), ("healpix", 32, 64, True),
         ("healpix", 40, 96, True), ("gl", None, 17, True), ("gl", None, 12, False)]


def _flms(rng, B, L, reality):
    return np.stack([signals.generate_flm(rng, L, 0, 0, reality) for _ in range(B)])


def _maps(rng, sampling, nside, L, reality, B=3):
    flm = torch.from_numpy(_flms(rng, B, L, reality)).cuda()
    f = sph.inverse_hip(flm, L, 0, nside, sampling, reality)
    if reality and f.is_complex():
        f = f.real.contiguous()
    return flm, f


@pytest.mark.parametrize("sampling,nside,L,reality", CASES)
def test_forward_is_rings_of_analysis(rng, sampling, nside, L, reality):
    flm, f = _maps(rng, sampling, nside, L, reality)
    plan = sph.get_plan(sampling, L, nside, 0, 0, reality, device=f.device)
    assert sph.spectra_supported(plan)
    G = sph.ring_analysis_hip(f, plan)
    assert torch.equal(sph.forward_rings_hip(G, L, plan), sph.forward_hip(f, L, 0, nside, sampling,
                                                                          reality, plan=plan))


@pytest.mark.parametrize("sampling,nside,L,reality", CASES)
def test_spectral_residual_matches_pixel_residual(rng, sampling, nside, L, reality):
    flm, f = _maps(rng, sampling, nside, L, reality)
    plan = sph.get_plan(sampling, L, nside, 0, 0, reality, device=f.device)
    x = flm + 0.01 * torch.from_numpy(_flms(rng, flm.shape[0], L, reality)).cuda()
    g = f + 0.1 * torch.randn_like(f)
    G = sph.ring_analysis_hip(g, plan)
    R = sph.ring_residual_hip(G, sph.inverse_rings_hip(x, plan), plan)
    real_out = reality and not g.is_complex()
    r_pix = sph.inverse_hip(x, L, 0, nside, sampling, reality, plan=plan, real_output=real_out,
                            sub_from=g)
    R_pix = sph.ring_analysis_hip(r_pix, plan)
    err = float((R - R_pix).abs().max() / G.abs().max())
    assert err < 1e-12, err
    # in place (R = G) gives the same
    G2 = G.clone()
    sph.ring_residual_hip(G2, sph.inverse_rings_hip(x, plan), plan, out=G2)
    assert torch.equal(G2, R)


def test_spectral_fp64_refinement_matches_pixel(rng):
    nside, L, B = 32, 64, 3
    flm, f = _maps(rng, "healpix", nside, L, True, B)
    f = f + 1e-3 * torch.randn_like(f)  # not band-limited: the iteration has work to do
    a = sph.forward_iterative_hip(f, L, 0, nside, "healpix", True, 0, 4, spectral=True)
    b = sph.forward_iterative_hip(f, L, 0, nside, "healpix", True, 0, 4, spectral=False)
    err = float((a - b).abs().max() / b.abs().max())
    assert err < 1e-12, err


@pytest.mark.parametrize("precision,refresh", [("mixed", 0), ("bf16", 3)])
def test_spectral_mixed_refinement_reaches_fp64_target(rng, precision, refresh):
    nside, L, B = 64, 128, 4
    flm_true = torch.from_numpy(_flms(rng, B, L, True)).cuda()
    f = sph.inverse_hip(flm_true, L, 0, nside, "healpix", True).real.contiguous()
    hist = []
    refine.refine_mixed(f, L, 0, nside, "healpix", True, 12, history=hist, flm_true=flm_true,
                        precision=precision, refresh_every=refresh, spectral=True)
    print(precision, ["%.1e" % e for e in hist])
    assert min(hist) <= 1e-7


def test_spectra_unsupported_for_resampled_forward():
    plan = sph.get_plan("mw", 8, None, 0, 0, True)
    assert not sph.spectra_supported(plan)
    with pytest.raises(ValueError):
        sph.ring_analysis_hip(torch.zeros((1, 8, 15), dtype=torch.float64, device="cuda"), plan)


@pytest.mark.parametrize("L", [1, 2, 7, 33])
def test_flm_pack_unpack_round_trip(rng, L):
    flm = torch.from_numpy(_flms(rng, 3, L, True)).cuda()
    p = sph.flm_pack_hip(flm, L)
    assert p.shape == (3, L * (L + 1) // 2)
    assert torch.equal(sph.flm_unpack_hip(p, L), flm)
    # packed order: (l, m) at m L - m (m - 1) / 2 + l - m
    pn, fn = p.cpu().numpy(), flm.cpu().numpy()
    for m in range(L):
        for l in range(m, L):
            assert pn[1, m * L - m * (m - 1) // 2 + l - m] == fn[1, l, L - 1 + m]


@pytest.mark.parametrize("precision", ["fp64", "bf16x1", "bf16", "mixed"])
def test_packed_rings_match_full_layout(rng, precision):
    """The latitudinal kernels give the same bits in the packed layout."""
    nside, L = 16, 32
    flm, f = _maps(rng, "healpix", nside, L, True, B=5)
    plan = sph.get_plan("healpix", L, nside, 0, 0, True, device=f.device)
    G = sph.ring_analysis_hip(f, plan)
    full = sph.forward_rings_hip(G, L, plan, precision)
    pk = sph.forward_rings_hip(G, L, plan, precision, packed=True)
    assert torch.equal(sph.flm_pack_hip(full, L), pk)
    assert torch.equal(sph.flm_unpack_hip(pk, L), full)
    Yf = sph.inverse_rings_hip(flm, plan, precision)
    Yp = sph.inverse_rings_hip(sph.flm_pack_hip(flm, L), plan, precision, packed=True)
    assert torch.equal(Yf, Yp)
    if precision in ("bf16", "bf16x1"):
        # against the FP64 synthesis: split ~1e-5, single BF16 ~4e-3 of max |Y|
        Y64 = sph.inverse_rings_hip(flm, plan)
        err = float((Yf - Y64).abs().max() / Y64.abs().max())
        assert err < (1e-4 if precision == "bf16" else 2e-2), err
    acc = pk.clone()
    sph.forward_rings_hip(G, L, plan, precision, accumulate_into=acc, packed=True)
    assert torch.allclose(acc, 2 * pk, rtol=1e-14, atol=0)


def test_packed_mlp_matches_full_layout(rng):
    from s2fft_amd.learned.precond import MLPPreconditioner
    L, B = 40, 3
    h = torch.from_numpy(_flms(rng, B, L, True)).cuda()
    pc = MLPPreconditioner.load(device="cuda")
    full = pc.apply(h, L, True, 0)
    pk = pc.apply(sph.flm_pack_hip(h, L), L, True, 0, packed=True)
    scale = float(full.abs().max())
    assert float((sph.flm_pack_hip(full, L) - pk).abs().max()) <= 1e-3 * scale
