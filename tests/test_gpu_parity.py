"""HIP transforms vs the CPU oracle on the same seeded inputs (needs a GPU).

Parameter grids follow the reference's tests/test_spherical_transform.py:14-21
(L in {6, 7}, spins {-2, 0, 1}, L_lower {0, 2}, nside {4, 5}) plus larger
sizes that exercise the Bluestein ring lengths (prime factors > 13) and the
high-m underflow handling.  Tolerance: the plain FP64 kernels must agree with
the oracle to 1e-12 norm-wise relative error.
"""

import numpy as np
import pytest
import torch

from oracle import ffts as offts
from oracle import pins
from oracle import pm as opm
from oracle import signals, transforms as otr
from oracle import geometry as ogeo
from tests.conftest import max_rel

import s2fft_amd
from s2fft_amd import _lib

pytestmark = pytest.mark.gpu
TOL = 1e-12


class _nullctx:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


def _flm(rng, L, spin=0, reality=False, L_lower=0):
    return signals.generate_flm(rng, L, L_lower, spin, reality)


@pytest.mark.parametrize("sampling", ["mw", "mwss", "dh", "gl"])
@pytest.mark.parametrize("L", [6, 7])
@pytest.mark.parametrize("spin", [-2, 0, 1])
@pytest.mark.parametrize("reality", [False, True])
@pytest.mark.parametrize("L_lower", [0, 2])
def test_inverse_equiangular(rng, sampling, L, spin, reality, L_lower):
    if reality and spin != 0:
        pytest.skip("reality only for spin 0")
    flm = _flm(rng, L, spin, reality, L_lower)
    ref = otr.inverse(flm, L, spin, sampling=sampling, reality=reality, L_lower=L_lower)
    got = s2fft_amd.inverse(flm, L, spin, sampling=sampling, method="hip", reality=reality,
                            L_lower=L_lower)
    assert got.shape == ref.shape
    assert max_rel(got, ref) < TOL


@pytest.mark.parametrize("sampling", ["mw", "mwss", "dh", "gl"])
@pytest.mark.parametrize("L", [6, 7, 9])
@pytest.mark.parametrize("spin", [-2, 0, 1])
@pytest.mark.parametrize("reality", [False, True])
@pytest.mark.parametrize("L_lower", [0, 2])
def test_forward_equiangular(rng, sampling, L, spin, reality, L_lower):
    if reality and spin != 0:
        pytest.skip("reality only for spin 0")
    flm = _flm(rng, L, spin, reality, L_lower)
    f = otr.inverse(flm, L, spin, sampling=sampling, reality=reality, L_lower=L_lower)
    ref = otr.forward(f, L, spin, sampling=sampling, reality=reality, L_lower=L_lower)
    got = s2fft_amd.forward(f, L, spin, sampling=sampling, method="hip", reality=reality,
                            L_lower=L_lower)
    assert max_rel(got, ref) < TOL
    assert max_rel(got, flm) < 1e-11   # sampling theorem: exact round trip


@pytest.mark.parametrize("nside", [4, 5, 16, 32])
@pytest.mark.parametrize("reality", [False, True])
def test_healpix_inverse_forward(rng, nside, reality):
    L = 2 * nside
    flm = _flm(rng, L, 0, reality)
    ref_f = otr.inverse(flm, L, 0, nside, "healpix", reality)
    got_f = s2fft_amd.inverse(flm, L, 0, nside, "healpix", "hip", reality)
    assert got_f.dtype == np.complex128
    assert max_rel(got_f, ref_f) < TOL
    f = np.real(ref_f) if reality else ref_f
    ref = otr.forward(f, L, 0, nside, "healpix", reality)
    got = s2fft_amd.forward(f, L, 0, nside, "healpix", "hip", reality)
    assert max_rel(got, ref) < TOL


@pytest.mark.parametrize("nside,L,spin", [(4, 8, 2), (5, 12, -1), (4, 9, 1), (8, 24, 0)])
def test_healpix_spin_and_oversampled(rng, nside, L, spin):
    """HEALPix with spin != 0: the reference's m-recursion drops NaN terms at
    some rings (see oracle.pm.dslices_exact), so the oracle takes its d-values
    from Wigner's explicit sum here; brute-force synthesis pins the inverse."""
    flm = _flm(rng, L, spin)
    with opm.exact_d():
        ref_f = otr.inverse(flm, L, spin, nside, "healpix")
        ref = otr.forward(ref_f, L, spin, nside, "healpix")
    got_f = s2fft_amd.inverse(flm, L, spin, nside, "healpix", "hip")
    assert max_rel(got_f, ref_f) < TOL
    pts = [pins.healpix_pix2ang(nside, p) for p in range(12 * nside * nside)]
    th = np.array([a for a, _ in pts])
    ph = np.array([b for _, b in pts])
    assert max_rel(got_f, pins.synthesis_bruteforce(flm, L, spin, th, ph)) < 1e-13
    got = s2fft_amd.forward(ref_f, L, spin, nside, "healpix", "hip")
    assert max_rel(got, ref) < TOL


@pytest.mark.parametrize("nside", [4, 5])
@pytest.mark.parametrize("n_iter", [0, 1, 3])
def test_healpix_iterative(rng, nside, n_iter):
    L = 2 * nside
    flm = _flm(rng, L, 0, True)
    f = np.real(otr.inverse(flm, L, 0, nside, "healpix", True))
    ref = otr.forward_iterative(f, L, 0, nside, "healpix", True, n_iter=n_iter)
    got = s2fft_amd.forward(f, L, 0, nside, "healpix", "hip", True, iter=n_iter)
    assert max_rel(got, ref) < TOL


@pytest.mark.parametrize("nside", [4, 5, 32])
@pytest.mark.parametrize("L_mult", [2, 3])
@pytest.mark.parametrize("reality", [False, True])
def test_healpix_ring_fft_entry(rng, nside, L_mult, reality):
    L = L_mult * nside
    npix = 12 * nside * nside
    f = rng.standard_normal(npix) + (0 if reality else 1j * rng.standard_normal(npix))
    f = f.astype(np.complex128)
    ref = offts.healpix_fft(f, L, nside, reality)
    ft = torch.from_numpy(f).cuda()
    ftm = torch.empty((4 * nside - 1, 2 * L), dtype=torch.complex128, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.lib().s2h_healpix_fft_prepare(nside, L))
    _lib.check(_lib.lib().s2h_healpix_fft(nside, L, int(reality), ft.data_ptr(), ftm.data_ptr(),
                                          1, s))
    assert max_rel(ftm.cpu().numpy(), ref) < TOL
    back_ref = offts.healpix_ifft(ref.copy(), L, nside, reality)
    back = torch.empty(npix, dtype=torch.complex128, device="cuda")
    _lib.check(_lib.lib().s2h_healpix_ifft(nside, L, int(reality), ftm.data_ptr(),
                                           back.data_ptr(), 1, s))
    assert max_rel(back.cpu().numpy(), back_ref) < TOL


@pytest.mark.parametrize("sampling,L,nside", [("healpix", 8, 4), ("dh", 7, None),
                                              ("gl", 6, None), ("mw", 7, None)])
@pytest.mark.parametrize("spin", [0, 2])
def test_latitudinal_entry(rng, sampling, L, nside, spin):
    """s2h_lat_inverse / s2h_lat_forward vs the oracle's latitudinal steps."""
    plan = _lib.Plan(sampling, L, nside or 0, spin, 0, False, _lib.NORM_NONE)
    beta = ogeo.thetas(L, sampling, nside)
    interior = slice(0, L - 1) if sampling == "mw" else slice(None)
    flm = _flm(rng, L, spin)
    # spin != 0: exact d-values (the m-recursion leaves junk in rows l < |spin|
    # and drops NaN terms on some HEALPix rings)
    ctx = opm.exact_d() if spin else _nullctx()
    with ctx:
        ref = opm.inverse_latitudinal_step(flm, beta, L, spin, nside, sampling)
    nrow, ncol = ref.shape
    d_flm = torch.from_numpy(flm).cuda()
    d_ftm = torch.empty((nrow, ncol), dtype=torch.complex128, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.lib().s2h_lat_inverse(plan.handle, d_flm.data_ptr(), d_ftm.data_ptr(), 1, s))
    got = d_ftm.cpu().numpy()
    assert max_rel(got[interior], ref[interior]) < TOL
    ftm = rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)
    if sampling == "mw":
        ftm[L - 1] = 0
    with (opm.exact_d() if spin else _nullctx()):
        ref2 = opm.forward_latitudinal_step(ftm, beta, L, spin, nside, sampling)
    d_in = torch.from_numpy(ftm).cuda()
    d_out = torch.empty((L, 2 * L - 1), dtype=torch.complex128, device="cuda")
    _lib.check(_lib.lib().s2h_lat_forward(plan.handle, d_in.data_ptr(), d_out.data_ptr(), 1, s))
    assert max_rel(d_out.cpu().numpy(), ref2) < TOL


def test_batched_equals_single(rng):
    """A batch gives each map's single-map result.  Real maps share their ring
    FFTs in pairs (z = x + i y), so a map's rounding depends on its partner:
    agreement to 1e-14 rather than bit for bit."""
    nside, L = 8, 16
    flms = np.stack([_flm(rng, L, 0, True) for _ in range(3)])
    f_b = s2fft_amd.inverse(flms, L, 0, nside, "healpix", "hip", True)
    for i in range(3):
        f_i = s2fft_amd.inverse(flms[i], L, 0, nside, "healpix", "hip", True)
        assert max_rel(f_b[i], f_i) < 1e-14
    g_b = s2fft_amd.forward(np.real(f_b), L, 0, nside, "healpix", "hip", True)
    for i in range(3):
        g_i = s2fft_amd.forward(np.real(f_b[i]), L, 0, nside, "healpix", "hip", True)
        assert max_rel(g_b[i], g_i) < 1e-14


@pytest.mark.parametrize("nside", [64, 128])
def test_healpix_larger_vs_oracle(rng, nside):
    """Larger HEALPix case: Bluestein rings and deep polar underflow.

    At nside 128 the reference's own m-recursion is 4.5e-12 (norm-wise) away
    from an 80-bit evaluation of the same sums, so agreement with the oracle
    is asserted within 1e-11 there and the latitudinal step is pinned to the
    long-double evaluation at 1e-12 by test_latitudinal_vs_longdouble."""
    L = 2 * nside
    tol = TOL if nside <= 64 else 1e-11
    flm = _flm(rng, L, 0, True)
    ref_f = otr.inverse(flm, L, 0, nside, "healpix", True)
    got_f = s2fft_amd.inverse(flm, L, 0, nside, "healpix", "hip", True)
    assert max_rel(got_f, ref_f) < tol
    f = np.real(ref_f)
    ref = otr.forward(f, L, 0, nside, "healpix", True)
    got = s2fft_amd.forward(f, L, 0, nside, "healpix", "hip", True)
    assert max_rel(got, ref) < tol


@pytest.mark.parametrize("nside", [128])
def test_latitudinal_vs_longdouble(rng, nside):
    """Spin-0 latitudinal synthesis vs the 80-bit pin at 1e-12."""
    L = 2 * nside
    flm = _flm(rng, L, 0, True)
    plan = _lib.Plan("healpix", L, nside, 0, 0, True, _lib.NORM_NONE)
    d_flm = torch.from_numpy(flm).cuda()
    d_ftm = torch.empty((4 * nside - 1, 2 * L), dtype=torch.complex128, device="cuda")
    _lib.check(_lib.lib().s2h_lat_inverse(plan.handle, d_flm.data_ptr(), d_ftm.data_ptr(), 1,
                                          torch.cuda.current_stream().cuda_stream))
    got = d_ftm.cpu().numpy()[:, L:]
    ref = pins.lat_inverse_longdouble(flm, ogeo.thetas(L, "healpix", nside), L)
    assert max_rel(got, ref) < 1e-12


@pytest.mark.parametrize("sampling,nside", [("mw", None), ("mwss", None), ("dh", None),
                                            ("gl", None), ("healpix", 4)])
@pytest.mark.parametrize("reality", [False, True])
@pytest.mark.parametrize("N", [1, 2, 3])
def test_wigner_vs_oracle(rng, sampling, nside, reality, N):
    """SO(3) transforms (reference tests/test_wigner_transform.py:37-79 grid)."""
    L = 8 if sampling == "healpix" else 6
    flmn = signals.generate_flmn(rng, L, N, reality=reality)
    ctx = opm.exact_d() if (sampling == "healpix" and N > 1) else _nullctx()
    with ctx:
        ref_f = otr.wigner_inverse(flmn, L, N, nside, sampling, reality)
        ref = otr.wigner_forward(ref_f, L, N, nside, sampling, reality)
    got_f = s2fft_amd.wigner.inverse(flmn, L, N, nside, sampling, "hip", reality)
    assert got_f.shape == ref_f.shape
    assert max_rel(got_f, ref_f) < TOL
    got = s2fft_amd.wigner.forward(ref_f, L, N, nside, sampling, "hip", reality)
    assert max_rel(got, ref) < TOL
    if sampling != "healpix":
        assert max_rel(got, flmn) < 1e-11


def test_mfma_f64_layout():
    """The A/B/result lane layouts the MFMA forward kernel assumes."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((16, 4))
    B = rng.standard_normal((4, 16))
    lanes = np.arange(64)
    a = A[lanes % 16, lanes // 16]          # A[i][k] at lane i + 16 k
    b = B[lanes // 16, lanes % 16]          # B[k][j] at lane j + 16 k
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = torch.zeros(256, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().s2h_probe_mfma_f64(da.data_ptr(), db.data_ptr(), out.data_ptr(),
                                             torch.cuda.current_stream().cuda_stream))
    got = out.cpu().numpy().reshape(64, 4)
    C = A @ B
    # FP64: register r of lane l holds row 4r + l//16
    expect = np.array([[C[4 * r + l // 16, l % 16] for r in range(4)] for l in range(64)])
    assert np.allclose(got, expect, atol=1e-13)


def test_mfma_f32_layout():
    """FP32 MFMA layouts the mixed-precision kernels assume."""
    rng = np.random.default_rng(6)
    A = rng.standard_normal((16, 4)).astype(np.float32)
    B = rng.standard_normal((4, 16)).astype(np.float32)
    lanes = np.arange(64)
    a = np.ascontiguousarray(A[lanes % 16, lanes // 16])
    b = np.ascontiguousarray(B[lanes // 16, lanes % 16])
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = torch.zeros(256, dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().s2h_probe_mfma_f32(da.data_ptr(), db.data_ptr(), out.data_ptr(),
                                             torch.cuda.current_stream().cuda_stream))
    got = out.cpu().numpy().reshape(64, 4).astype(np.float64)
    C = A.astype(np.float64) @ B.astype(np.float64)
    # FP32: register r of lane l holds row 4 (l//16) + r
    expect = np.array([[C[4 * (l // 16) + r, l % 16] for r in range(4)] for l in range(64)])
    assert np.allclose(got, expect, atol=1e-5)


@pytest.mark.parametrize("sampling,nside,L,spin,reality,B", [
    ("healpix", 16, 32, 0, True, 17), ("healpix", 8, 16, 0, False, 4),
    ("gl", None, 12, 2, False, 5), ("mwss", None, 10, 0, True, 8), ("mw", None, 9, 1, False, 6)])
def test_batched_mma_inverse_matches_oracle(rng, sampling, nside, L, spin, reality, B):
    flms = np.stack([_flm(rng, L, spin, reality) for _ in range(B)])
    got = s2fft_amd.inverse(flms, L, spin, nside, sampling, "hip", reality)
    for i in range(B):
        ref = otr.inverse(flms[i], L, spin, nside, sampling, reality)
        assert max_rel(got[i], ref) < TOL


@pytest.mark.parametrize("sampling,nside,L,spin,reality,B", [
    ("healpix", 16, 32, 0, True, 9), ("healpix", 8, 16, 0, False, 4),
    ("gl", None, 12, 2, False, 5), ("dh", None, 10, 0, True, 8), ("mw", None, 9, 0, True, 4)])
def test_batched_mfma_forward_matches_oracle(rng, sampling, nside, L, spin, reality, B):
    flms = np.stack([_flm(rng, L, spin, reality) for _ in range(B)])
    fs = np.stack([otr.inverse(flms[i], L, spin, nside, sampling, reality) for i in range(B)])
    if reality:
        fs = np.real(fs)
    got = s2fft_amd.forward(fs, L, spin, nside, sampling, "hip", reality)
    for i in range(B):
        ref = otr.forward(fs[i], L, spin, nside, sampling, reality)
        assert max_rel(got[i], ref) < TOL


@pytest.mark.parametrize("sampling", ["mw", "mwss"])
@pytest.mark.parametrize("L", [6, 7])
def test_forward_reality_non_bandlimited(rng, sampling, L):
    """A real map that is not band-limited: the reference resamples the complex
    map and keeps its real part before the rfft; the HIP path must agree."""
    f = rng.standard_normal(ogeo.f_shape(L, sampling))
    ref = otr.forward(f, L, 0, sampling=sampling, reality=True)
    got = s2fft_amd.forward(f, L, 0, sampling=sampling, method="hip", reality=True)
    assert max_rel(got, ref) < TOL
    refc = otr.forward(f + 0j, L, 0, sampling=sampling, reality=False)
    gotc = s2fft_amd.forward(f + 0j, L, 0, sampling=sampling, method="hip", reality=False)
    assert max_rel(gotc, refc) < TOL


@pytest.mark.parametrize("sampling,nside,L,reality", [("healpix", 8, 16, True), ("mw", None, 9, True),
                                                      ("gl", None, 8, False)])
def test_fused_residual_and_accumulate(rng, sampling, nside, L, reality):
    """inverse(sub_from=...) == f - inverse, real_output == real part,
    forward(accumulate_into=...) == flm + forward: the refinement's fused passes."""
    from s2fft_amd.transforms import spherical as sph
    B = 3
    flm = torch.from_numpy(np.stack([_flm(rng, L, 0, reality) for _ in range(B)])).cuda()
    f = sph.inverse_hip(flm, L, 0, nside, sampling, reality)
    if reality:
        fr = f.real.contiguous() if f.is_complex() else f
        ro = sph.inverse_hip(flm, L, 0, nside, sampling, reality, real_output=True)
        assert ro.dtype == torch.float64
        assert max_rel(ro.cpu().numpy(), fr.cpu().numpy()) < 1e-15
    else:
        fr = f
    g = fr + 0.25 * torch.randn_like(fr)
    r = sph.inverse_hip(flm, L, 0, nside, sampling, reality, real_output=reality, sub_from=g)
    # the real-output path pairs maps in one complex FFT: rounding at the
    # level of |f|, so the error is measured against |g|, not the small |r|
    err = (r - (g - fr)).abs().max() / g.abs().max()
    assert float(err) < 1e-14
    base = sph.forward_hip(g, L, 0, nside, sampling, reality)
    acc = base.clone()
    sph.forward_hip(r, L, 0, nside, sampling, reality, accumulate_into=acc)
    ref = base + sph.forward_hip(r, L, 0, nside, sampling, reality)
    assert max_rel(acc.cpu().numpy(), ref.cpu().numpy()) < 1e-14
