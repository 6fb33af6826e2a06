"""Tiled ring spectra (blocks of 8 rows, S2H_SPECTRA_TILED) against the
ring-major layout (needs a GPU).

The latitudinal kernels only change their addresses with the layout, so
their results agree bit for bit after re-tiling.  The tiled alias kernel
forms the same per-ring sums; residues of rings shorter than 256 and the
dot's partial sums run in another fixed order, so those agree to 1e-13
relative.  Sizes: HEALPix nside 16 and 64 (63 and 255 rings: a partial last
block), one map (vector-ALU kernels) and 16 maps (matrix-core kernels).
"""

import numpy as np
import pytest
import torch

from s2fft_amd.transforms import spherical as sph

pytestmark = pytest.mark.gpu


def _to_tiled(G, ncol):
    B, n = G.shape
    return G.view(B, n // ncol // 8, 8, ncol).transpose(2, 3).contiguous().view(B, n)


def _from_tiled(T, ncol):
    B, n = T.shape
    return T.view(B, n // ncol // 8, ncol, 8).transpose(2, 3).contiguous().view(B, n)


def _rows(nside):
    return 4 * nside - 1


def _valid(X, nside, L):
    B = X.shape[0]
    return X.view(B, -1, L)[:, :_rows(nside)]


def _random_spectra(plan, B, nside, L, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    G = sph._new_spectra(plan, B, torch.device("cuda"))
    G.zero_()
    v = torch.randn((B, _rows(nside), L, 2), generator=g, device="cuda", dtype=torch.float64)
    G.view(B, -1, L)[:, :_rows(nside)] = torch.view_as_complex(v)
    return G


def _flm(B, L, seed):
    from s2fft_amd.utils.signal_generator import generate_flm
    flm = np.stack([generate_flm(np.random.default_rng(seed + i), L, reality=True)
                    for i in range(B)])
    return sph.flm_pack_hip(torch.from_numpy(flm).cuda(), L)


@pytest.mark.parametrize("nside", [16, 64])
@pytest.mark.parametrize("B", [1, 16])
@pytest.mark.parametrize("precision", ["fp64", "mixed", "bf16", "bf16x1"])
def test_inverse_rings_tiled_bitwise(nside, B, precision):
    L = 2 * nside
    plan = sph.get_plan("healpix", L, nside, 0, 0, True)
    x = _flm(B, L, 40 + nside)
    Yr = sph.inverse_rings_hip(x, plan, precision, packed=True)
    Yt = sph.inverse_rings_hip(x, plan, precision, packed=True, tiled=True)
    assert torch.equal(_valid(_from_tiled(Yt, L), nside, L), _valid(Yr, nside, L))


@pytest.mark.parametrize("nside", [16, 64])
@pytest.mark.parametrize("B", [1, 16])
@pytest.mark.parametrize("precision", ["fp64", "mixed", "bf16", "bf16x1"])
def test_forward_rings_tiled_bitwise(nside, B, precision):
    L = 2 * nside
    plan = sph.get_plan("healpix", L, nside, 0, 0, True)
    G = _random_spectra(plan, B, nside, L, 7 + B)
    fr = sph.forward_rings_hip(G, L, plan, precision, packed=True)
    ft = sph.forward_rings_hip(_to_tiled(G, L), L, plan, precision, packed=True, tiled=True)
    assert torch.equal(fr, ft)
    # the reference flm layout as well
    fr = sph.forward_rings_hip(G, L, plan, precision)
    ft = sph.forward_rings_hip(_to_tiled(G, L), L, plan, precision, tiled=True)
    assert torch.equal(fr, ft)


def _close(a, b, rel):
    scale = b.abs().max().item()
    return (a - b).abs().max().item() <= rel * scale


@pytest.mark.parametrize("nside", [16, 64])
@pytest.mark.parametrize("B", [1, 3, 16])
def test_ring_residual_tiled(nside, B):
    L = 2 * nside
    plan = sph.get_plan("healpix", L, nside, 0, 0, True)
    G = _random_spectra(plan, B, nside, L, 11)
    Y = _random_spectra(plan, B, nside, L, 12)
    alpha = torch.linspace(0.5, 1.5, B, device="cuda", dtype=torch.float64)
    Rr, sr = sph.ring_residual_hip(G, Y, plan, alpha=alpha, sumsq=True)
    Yt = _to_tiled(Y, L)
    # G ring-major in, R tiled out
    Rt, st = sph.ring_residual_hip(G, Yt, plan, alpha=alpha, sumsq=True, tiled=True,
                                   g_tiled=False)
    assert _close(_valid(_from_tiled(Rt, L), nside, L), _valid(Rr, nside, L), 1e-13)
    assert torch.allclose(st, sr, rtol=1e-12, atol=0)
    # G tiled, in place
    Gt = _to_tiled(G, L)
    sph.ring_residual_hip(Gt, Yt, plan, out=Gt, alpha=alpha, tiled=True)
    assert _close(_valid(_from_tiled(Gt, L), nside, L), _valid(Rr, nside, L), 1e-13)


@pytest.mark.parametrize("nside", [16, 64])
@pytest.mark.parametrize("B", [1, 16])
def test_ring_dot_tiled(nside, B):
    L = 2 * nside
    plan = sph.get_plan("healpix", L, nside, 0, 0, True)
    # a synthesised Y: the dot is the curvature <p, F A I p> > 0
    x = _flm(B, L, 90)
    Yr = sph.inverse_rings_hip(x, plan, "fp64", packed=True)
    Yt = sph.inverse_rings_hip(x, plan, "fp64", packed=True, tiled=True)
    num = torch.ones(B, device="cuda", dtype=torch.float64)
    pr, ar = sph.ring_dot_hip(Yr, plan, num=num)
    pt, at = sph.ring_dot_hip(Yt, plan, num=num, tiled=True)
    assert torch.all(pr > 0)
    assert torch.allclose(pt, pr, rtol=1e-12, atol=0)
    assert torch.allclose(at, ar, rtol=1e-12, atol=0)


@pytest.mark.parametrize("nside", [16, 64])
@pytest.mark.parametrize("B", [1, 5, 16])
@pytest.mark.parametrize("precision", ["fp64", "mixed", "bf16", "bf16x1"])
def test_inverse_rings_dot_fused(nside, B, precision):
    """The synthesis with the dot fused over the rows without aliasing: the
    same spectra bit for bit, and the dot of the separate pass to 1e-12."""
    L = 2 * nside
    plan = sph.get_plan("healpix", L, nside, 0, 0, True)
    x = _flm(B, L, 130 + B)
    num = torch.linspace(1.0, 2.0, B, device="cuda", dtype=torch.float64)
    Yt = sph.inverse_rings_hip(x, plan, precision, packed=True, tiled=True)
    pr, ar = sph.ring_dot_hip(Yt, plan, num=num, tiled=True)
    Yf, pf, af = sph.inverse_rings_dot_hip(x, plan, precision, packed=True, num=num)
    assert torch.equal(_valid(Yf, nside, L), _valid(Yt, nside, L))
    assert torch.allclose(pf, pr, rtol=1e-12, atol=0)
    assert torch.allclose(af, ar, rtol=1e-12, atol=0)
    Y2, p2 = sph.inverse_rings_dot_hip(x, plan, precision, packed=True)
    assert torch.equal(p2, pf)


def test_tiled_refuses_complex_maps():
    nside, L = 16, 32
    plan = sph.get_plan("healpix", L, nside, 0, 0, False)
    Y = sph._new_spectra(plan, 1, torch.device("cuda"))
    with pytest.raises(NotImplementedError):
        sph.ring_dot_hip(Y, plan, tiled=True)
    with pytest.raises(NotImplementedError):
        sph.ring_residual_hip(Y.clone(), Y, plan, tiled=True)
