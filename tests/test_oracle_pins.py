"""Pins of the CPU oracle by independent mathematics (CPU only).

The reference ships no golden vectors for this path (its cached test data is
not in the repository and its external oracles -- pyssht, healpy -- are not
installed), so every oracle function is pinned here by something that does
not share its code:
  * Wigner d^l_{mn} from Wigner's explicit sum in multiprecision (pins the
    Price-McEwen recursion restatement, signs and spin convention);
  * scipy.special.sph_harm_y for spin 0 (pins geometry, phases, FFT
    normalisation and the m-column layout of the inverse);
  * the HEALPix pix2ang closed form (pins ring colatitudes and azimuths);
  * sampling theorems (pins every forward on MW/MWSS/DH/GL), the HEALPix
    adjoint identity (pins the HEALPix forward), refinement convergence;
  * numpy.polynomial.legendre.leggauss and closed-form integrals (weights);
  * brute-force DFTs (ring FFTs, periodic extension and folding).
Each check is chosen so that a plausible mistake -- a dropped sign, a wrong m
offset, a transposed operand, a missing weight -- fails it.
"""

import numpy as np
import pytest

from oracle import ffts, geometry as geo, pins, pm, signals, transforms as T
from tests.conftest import max_rel


# ------------------------------------------------------------ Wigner d values
@pytest.mark.parametrize("spin", [-2, -1, 0, 1, 3])
@pytest.mark.parametrize("reality", [False, True])
def test_pm_dslices_match_explicit_wigner_d(spin, reality):
    if reality and spin != 0:
        pytest.skip("reality only for spin 0")
    L = 9
    beta = np.array([0.03, 0.4, 1.1, np.pi / 2, 2.0, 2.9, 3.11])
    n = -spin
    for m, D in pm.dslices(beta, L, spin, reality):
        for l in range(max(abs(m), abs(spin)), L):
            exact = np.array([pins.wigner_d(l, m, n, b) for b in beta])
            got = np.nan_to_num(D[:, l])
            # the m-recursion loses digits next to the poles as |spin| grows
            # (1.7e-11 absolute at theta = 3.11, spin 3); sign or index slips
            # give O(0.1) errors
            tol = 1e-12 if abs(spin) < 3 else 1e-10
            assert np.max(np.abs(got - exact)) < tol, (m, l)


def test_pm_dslices_L_lower_consistent():
    L, spin = 8, 1
    beta = geo.thetas(L, "dh")
    full = {m: D for m, D in pm.dslices(beta, L, spin)}
    for m, D in pm.dslices(beta, L, spin, L_lower=3):
        lo = max(abs(m), abs(spin), 3)
        assert np.allclose(np.nan_to_num(D[:, lo - 3:]), np.nan_to_num(full[m][:, lo:]),
                           atol=1e-13)


# ------------------------------------------------------------ full transforms
@pytest.mark.parametrize("sampling", ["mw", "mwss", "dh", "gl"])
@pytest.mark.parametrize("reality", [False, True])
def test_inverse_spin0_matches_scipy(rng, sampling, reality):
    L = 6
    flm = signals.generate_flm(rng, L, reality=reality)
    f = T.inverse(flm, L, 0, sampling=sampling, reality=reality)
    th = geo.thetas(L, sampling)
    nphi = geo.nphi_equiang(L, sampling)
    ph = 2 * np.pi * np.arange(nphi) / nphi
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    brute = pins.synthesis_bruteforce(flm, L, 0, TH, PH, use_scipy=True)
    assert max_rel(f, brute) < 1e-13


@pytest.mark.parametrize("sampling", ["mw", "mwss", "dh", "gl"])
@pytest.mark.parametrize("spin", [-2, 1, 2])
def test_inverse_spin_matches_explicit_sylm(rng, sampling, spin):
    L = 5
    flm = signals.generate_flm(rng, L, spin=spin)
    f = T.inverse(flm, L, spin, sampling=sampling)
    th = geo.thetas(L, sampling)
    nphi = geo.nphi_equiang(L, sampling)
    ph = 2 * np.pi * np.arange(nphi) / nphi
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    brute = pins.synthesis_bruteforce(flm, L, spin, TH, PH)
    assert max_rel(f, brute) < 1e-13


def test_sylm_spin0_equals_scipy():
    th = np.array([0.1, 0.9, 2.3])
    ph = np.array([0.3, 4.0, 1.7])
    for l in range(5):
        for m in range(-l, l + 1):
            assert np.allclose(pins.sylm(l, m, 0, th, ph), pins.ylm_scipy(l, m, th, ph),
                               atol=1e-14)


@pytest.mark.parametrize("nside", [1, 2, 3, 4, 5, 8])
def test_healpix_geometry_matches_pix2ang(nside):
    offs, nph = ffts.ring_offsets(nside)
    th = geo.thetas(2 * nside, "healpix", nside)
    for t, (o, n) in enumerate(zip(offs, nph)):
        phis = geo.healpix_ring_phis(t, nside)
        assert np.isclose(geo.healpix_phi0(np.array([t]), nside)[0], phis[0], atol=1e-15)
        for p in range(n):
            a, b = pins.healpix_pix2ang(nside, o + p)
            assert abs(a - th[t]) < 1e-12 and abs(b - phis[p]) < 1e-12


@pytest.mark.parametrize("nside", [2, 4, 5])
@pytest.mark.parametrize("reality", [False, True])
def test_healpix_inverse_matches_scipy(rng, nside, reality):
    L = 2 * nside
    flm = signals.generate_flm(rng, L, reality=reality)
    f = T.inverse(flm, L, 0, nside, "healpix", reality)
    pts = [pins.healpix_pix2ang(nside, p) for p in range(12 * nside * nside)]
    th = np.array([a for a, _ in pts])
    ph = np.array([b for _, b in pts])
    brute = pins.synthesis_bruteforce(flm, L, 0, th, ph, use_scipy=True)
    assert max_rel(f, brute) < 1e-13


@pytest.mark.parametrize("nside,L,spin", [(4, 8, 0), (4, 9, 1), (3, 7, -2)])
def test_healpix_adjoint_identity(rng, nside, L, spin):
    """<I flm, g> = (Npix / 4pi) <flm, F g> pins the HEALPix forward."""
    npix = 12 * nside * nside
    flm = signals.generate_flm(rng, L, spin=spin)
    g = rng.standard_normal(npix) + 1j * rng.standard_normal(npix)
    lhs = np.vdot(T.inverse(flm, L, spin, nside, "healpix"), g)
    rhs = npix / (4 * np.pi) * np.vdot(flm, T.forward(g, L, spin, nside, "healpix"))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


@pytest.mark.parametrize("sampling", ["mw", "mwss", "dh", "gl"])
@pytest.mark.parametrize("spin", [-2, 0, 1])
@pytest.mark.parametrize("L_lower", [0, 2])
@pytest.mark.parametrize("reality", [False, True])
def test_sampling_theorem_round_trip(rng, sampling, spin, L_lower, reality):
    if reality and spin != 0:
        pytest.skip("reality only for spin 0")
    L = 7
    flm = signals.generate_flm(rng, L, L_lower, spin, reality)
    f = T.inverse(flm, L, spin, sampling=sampling, reality=reality, L_lower=L_lower)
    back = T.forward(f, L, spin, sampling=sampling, reality=reality, L_lower=L_lower)
    assert max_rel(back, flm) < 1e-12


def test_healpix_refinement_converges(rng):
    nside, L = 4, 8
    flm = signals.generate_flm(rng, L, reality=True)
    f = np.real(T.inverse(flm, L, 0, nside, "healpix", True))
    errs = [max_rel(T.forward_iterative(f, L, 0, nside, "healpix", True, n_iter=k), flm)
            for k in (0, 2, 4, 10)]
    assert all(a > b for a, b in zip(errs, errs[1:]))
    assert errs[-1] < 1e-9


# ------------------------------------------------------------ quadrature
@pytest.mark.parametrize("sampling", ["mw", "mwss", "dh", "gl"])
def test_quadrature_integral_closed_form(rng, sampling):
    """sum_t,p w_t f = sqrt(4 pi) f_00  (reference tests/test_quadrature.py:15-40)."""
    L = 6
    flm = signals.generate_flm(rng, L)
    f = T.inverse(flm, L, 0, sampling=sampling)
    q = geo.quad_weights(L, sampling)
    assert abs(np.sum(q[:, None] * f) - np.sqrt(4 * np.pi) * flm[0, L - 1]) < 1e-13


def test_gl_weights_match_leggauss():
    for L in (4, 9, 32):
        x, w = np.polynomial.legendre.leggauss(L)
        assert np.allclose(geo.quad_weights_gl(L), np.flip(w) * 2 * np.pi / (2 * L - 1),
                           atol=1e-14)


def test_healpix_weights_and_ring_counts():
    for nside in (1, 3, 16):
        offs, nph = ffts.ring_offsets(nside)
        assert nph.sum() == 12 * nside * nside
        assert np.allclose(geo.quad_weights_hp(nside).sum() * 0 + geo.quad_weights_hp(nside)[0],
                           4 * np.pi / (12 * nside * nside))


# ------------------------------------------------------------ ring FFTs
@pytest.mark.parametrize("nside,L", [(2, 4), (3, 7), (4, 8)])
def test_healpix_fft_matches_bruteforce(rng, nside, L):
    """ftm[t, m] = sum_p f_p exp(-i m 2 pi p / nphi): FFT + periodic extension."""
    npix = 12 * nside * nside
    f = rng.standard_normal(npix) + 1j * rng.standard_normal(npix)
    ftm = ffts.healpix_fft(f, L, nside)
    offs, nph = ffts.ring_offsets(nside)
    m = np.arange(-L, L)
    for t, (o, n) in enumerate(zip(offs, nph)):
        p = np.arange(n)
        brute = np.exp(-2j * np.pi * np.outer(m, p) / n) @ f[o : o + n]
        assert np.allclose(ftm[t], brute, atol=1e-12)


@pytest.mark.parametrize("nside,L", [(2, 4), (3, 7)])
def test_healpix_ifft_matches_bruteforce(rng, nside, L):
    """f_p = sum_m ftm[t, m] exp(+i m 2 pi p / nphi): folding + iFFT."""
    ftm = rng.standard_normal((4 * nside - 1, 2 * L)) + 1j * rng.standard_normal((4 * nside - 1, 2 * L))
    f = ffts.healpix_ifft(ftm.copy(), L, nside)
    offs, nph = ffts.ring_offsets(nside)
    m = np.arange(-L, L)
    for t, (o, n) in enumerate(zip(offs, nph)):
        p = np.arange(n)
        brute = np.exp(2j * np.pi * np.outer(p, m) / n) @ ftm[t]
        assert np.allclose(f[o : o + n], brute, atol=1e-12)


def test_dft_pin():
    x = np.random.default_rng(1).standard_normal(12) + 0j
    assert np.allclose(pins.dft(x), np.fft.fft(x))


# ------------------------------------------------------------ resampling
@pytest.mark.parametrize("spin", [0, 2])
def test_mw_to_mwss_is_exact_resampling(rng, spin):
    """MW -> MWSS resampling equals direct synthesis on the MWSS grid."""
    L = 6
    flm = signals.generate_flm(rng, L, spin=spin)
    f_mw = T.inverse(flm, L, spin, sampling="mw")
    f_ss = T.inverse(flm, L, spin, sampling="mwss")
    assert max_rel(ffts.mw_to_mwss(f_mw, L, spin), f_ss) < 1e-12
    up = ffts.upsample_by_two_mwss(f_ss, L, spin)
    th = geo.thetas(2 * L, "mwss")
    ph = 2 * np.pi * np.arange(2 * L) / (2 * L)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    assert max_rel(up, pins.synthesis_bruteforce(flm, L, spin, TH, PH)) < 1e-12


# ------------------------------------------------------------ signals
def test_generate_flm_draw_order_and_symmetry():
    rng = np.random.default_rng(0)
    L = 5
    flm = signals.generate_flm(rng, L, reality=True)
    rng2 = np.random.default_rng(0)
    m0 = rng2.standard_normal(L)
    assert np.array_equal(flm[:, L - 1].real, m0)
    for l in range(L):
        for m in range(1, l + 1):
            assert flm[l, L - 1 - m] == (-1) ** m * np.conj(flm[l, L - 1 + m])
        assert np.all(flm[l, L - 1 + l + 1 :] == 0)


def test_longdouble_pin_matches_explicit_sum(rng):
    """The 80-bit latitudinal pin agrees with Wigner's explicit sum."""
    L, nside = 8, 4
    beta = geo.thetas(L, "healpix", nside)
    flm = signals.generate_flm(rng, L, reality=True)
    got = pins.lat_inverse_longdouble(flm, beta, L)
    for m in range(L):
        exact = sum(flm[l, L - 1 + m] * np.array([pins.wigner_d(l, m, 0, b) for b in beta])
                    for l in range(m, L))
        assert np.max(np.abs(got[:, m] - exact)) < 1e-14


def test_oracle_pm_precision_at_nside_128():
    """Records the reference recursion's own error at L=256 against the 80-bit
    pin (4.5e-12 norm-wise on this input); the HIP kernel is held to 1e-12 of
    the pin instead (tests/test_gpu_parity.py)."""
    nside, L = 128, 256
    rng = np.random.default_rng(7)
    flm = signals.generate_flm(rng, L, reality=True)
    beta = geo.thetas(L, "healpix", nside)
    ref = pins.lat_inverse_longdouble(flm, beta, L)
    ora = pm.inverse_latitudinal_step(flm, beta, L, 0, nside, "healpix", reality=True)[:, L:]
    err = max_rel(ora, ref)
    assert 1e-13 < err < 1e-10


# ------------------------------------------------------------ Wigner (SO(3))
def _wigner_bruteforce(flmn, L, N, thetas, phis_per_ring, gammas):
    """f[g, point] = sum_{n,l,m} (2l+1)/(8 pi^2) flmn[n, l, m] conj(D^l_{mn}(alpha, beta, gamma))
    with D^l_{mn} = e^{-i m alpha} d^l_{mn}(beta) e^{-i n gamma}, d from Wigner's explicit sum
    (pins.wigner_d); points are (thetas[t], phis_per_ring[t][p]) flattened ring by ring."""
    G = len(gammas)
    cols = []
    for t, beta in enumerate(thetas):
        alpha = np.asarray(phis_per_ring[t], dtype=np.float64)
        ring = np.zeros((G, len(alpha)), dtype=np.complex128)
        for n in range(-N + 1, N):
            en = np.exp(1j * n * np.asarray(gammas))
            for l in range(abs(n), L):
                c = (2 * l + 1) / (8 * np.pi**2)
                for m in range(-l, l + 1):
                    v = flmn[N - 1 + n, l, L - 1 + m]
                    if v == 0:
                        continue
                    d = pins.wigner_d(l, m, n, beta)
                    ring += c * v * d * np.outer(en, np.exp(1j * m * alpha))
        cols.append(ring)
    return np.concatenate(cols, axis=1)


@pytest.mark.parametrize("sampling", ["mw", "mwss"])
@pytest.mark.parametrize("N", [1, 2, 3])
@pytest.mark.parametrize("reality", [False, True])
def test_wigner_inverse_and_forward_match_bruteforce_D(sampling, N, reality):
    """The oracle's SO(3) inverse against brute-force D^l_mn synthesis (1e-13),
    and its forward recovering flmn from the brute-force map (the MW / MWSS
    sampling theorem is exact, 1e-12).  A wrong sqrt((2l+1)/(16 pi^3)) or a
    (-1)^n applied in both directions fails the first check."""
    L = 6
    flmn = signals.generate_flmn(np.random.default_rng(40 + N), L, N, reality=reality)
    th = geo.thetas(L, sampling)
    nphi = geo.nphi_equiang(L, sampling)
    phis = [2 * np.pi * np.arange(nphi) / nphi] * len(th)
    gam = 2 * np.pi * np.arange(2 * N - 1) / (2 * N - 1)
    bf = _wigner_bruteforce(flmn, L, N, th, phis, gam).reshape(2 * N - 1, len(th), nphi)
    got = T.wigner_inverse(flmn, L, N, sampling=sampling, reality=reality)
    assert got.shape == bf.shape
    assert max_rel(got, bf) < 1e-13
    if reality:
        assert np.max(np.abs(bf.imag)) < 1e-13 * np.max(np.abs(bf))
        bf = bf.real
    back = T.wigner_forward(bf, L, N, sampling=sampling, reality=reality)
    assert max_rel(back, flmn) < 1e-12


@pytest.mark.parametrize("N", [1, 2])
def test_wigner_healpix_inverse_matches_bruteforce_D(N):
    nside, L = 2, 5
    flmn = signals.generate_flmn(np.random.default_rng(50 + N), L, N)
    th = geo.thetas(L, "healpix", nside)
    phis = [geo.healpix_ring_phis(t, nside) for t in range(4 * nside - 1)]
    gam = 2 * np.pi * np.arange(2 * N - 1) / (2 * N - 1)
    bf = _wigner_bruteforce(flmn, L, N, th, phis, gam)
    # the reference's m-recursion drops the terms of rings where one of its
    # intermediates is exactly zero (cos theta = +-1/3 at nside 2, spin != 0);
    # with the explicit-sum d the oracle is the exact synthesis
    with pm.exact_d():
        got = T.wigner_inverse(flmn, L, N, nside, "healpix")
    assert max_rel(got, bf) < 1e-13
    if N > 1:
        ref_like = T.wigner_inverse(flmn, L, N, nside, "healpix")
        bad = np.abs(ref_like - bf).max(axis=0) > 1e-10
        z = np.cos(np.concatenate([np.full(len(p), t) for t, p in zip(th, phis)]))
        assert bad.any() and np.allclose(np.abs(z[bad]), 1 / 3)


def test_generate_flmn_draw_order_and_symmetry():
    L, N = 5, 3
    flmn = signals.generate_flmn(np.random.default_rng(3), L, N, reality=True)
    r = np.random.default_rng(3)
    # the first draws are the n = -N+1 plane's m > 0 coefficients (tril order)
    el, m = np.tril_indices(L, k=-1, m=L)
    keep = el >= N - 1
    first = r.standard_normal(keep.sum())
    assert np.allclose(flmn[0, el[keep], L - 1 + m[keep] + 1].real, first)
    for n in range(-N + 1, N):
        for l in range(L):
            for mm in range(-l, l + 1):
                assert flmn[N - 1 - n, l, L - 1 - mm] == pytest.approx(
                    (-1) ** (mm + n) * np.conj(flmn[N - 1 + n, l, L - 1 + mm]))
            if l < abs(n):
                assert np.all(flmn[N - 1 + n, l] == 0)


@pytest.mark.parametrize("sampling,nside,L,rings", [
    ("healpix", 4, 9, [0, 3, 5, 7, 12, 14]), ("mw", None, 7, [0, 2, 5]),
    ("mwss", None, 6, [1, 3, 5]), ("gl", None, 6, [0, 5]), ("dh", None, 5, [0, 4, 9])])
@pytest.mark.parametrize("reality", [False, True])
@pytest.mark.parametrize("spin", [0, 1])
def test_inverse_rings_is_a_ring_subset_of_inverse(sampling, nside, L, rings, reality, spin):
    if reality and spin:
        pytest.skip("reality only for spin 0")
    flm = signals.generate_flm(np.random.default_rng(5), L, 0, spin, reality)
    full = np.asarray(T.inverse(flm, L, spin, nside, sampling, reality))
    if sampling == "healpix":
        offs, nph = ffts.ring_offsets(nside)
        want = np.concatenate([full[offs[t]:offs[t] + nph[t]] for t in rings])
    else:
        want = full[rings].reshape(-1)
    got = T.inverse_rings(flm, L, rings, spin, nside, sampling, reality)
    assert max_rel(got, want) < 1e-14


# ------------------------------------------------------------ precompute path
from oracle import precompute as opc  # noqa: E402


@pytest.mark.parametrize("spin", [0, 1, -2])
def test_compute_all_slices_matches_explicit_wigner_d(spin):
    L = 7
    beta = np.array([0.2, 1.0, np.pi / 2, 2.5, 3.0])
    dl = opc.compute_all_slices(beta, L, spin)
    assert dl.shape == (2 * L - 1, len(beta), L)
    for m in range(-(L - 1), L):
        for l in range(L):
            exact = [pins.wigner_d(l, m, -spin, b) if l >= max(abs(m), abs(spin)) else 0.0
                     for b in beta]
            assert np.max(np.abs(dl[L - 1 + m, :, l] - exact)) < 1e-12, (m, l)


@pytest.mark.parametrize("sampling,nside,L,spin,reality", [
    ("mw", None, 6, 0, True), ("mwss", None, 6, 1, False), ("gl", None, 5, 0, False),
    ("dh", None, 5, -1, False), ("healpix", 2, 5, 0, True), ("healpix", 2, 5, 0, False)])
@pytest.mark.parametrize("forward", [False, True])
def test_spin_kernel_and_transforms_match_the_on_the_fly_oracle(sampling, nside, L, spin,
                                                                reality, forward):
    """The kernel from the m-recursion equals the explicit-sum one, and the
    precompute transforms equal the on-the-fly transforms (the reference's
    precompute tests compare the two paths, test_spherical_precompute.py:57-58)."""
    k = opc.spin_spherical_kernel(L, spin, reality, sampling, nside, forward)
    ke = opc.spin_spherical_kernel(L, spin, reality, sampling, nside, forward, exact=True)
    assert np.max(np.abs(k - ke)) < 1e-12 * np.max(np.abs(ke))
    flm = signals.generate_flm(np.random.default_rng(9), L, 0, spin, reality)
    f = T.inverse(flm, L, spin, nside, sampling, reality)
    if not forward:
        got = opc.inverse_transform(flm, ke, L, sampling, reality, spin, nside)
        assert max_rel(got, f) < 1e-12
    else:
        fin = np.real(f) if reality else f
        got = opc.forward_transform(fin, ke, L, sampling, reality, spin, nside)
        assert max_rel(got, T.forward(fin, L, spin, nside, sampling, reality)) < 1e-12


@pytest.mark.parametrize("sampling,nside,L", [("mw", None, 5), ("healpix", 2, 5)])
@pytest.mark.parametrize("forward", [False, True])
def test_wigner_kernel_m_recursion_matches_explicit_sum(sampling, nside, L, forward):
    k = opc.wigner_kernel(L, 2, False, sampling, nside, forward)
    ke = opc.wigner_kernel(L, 2, False, sampling, nside, forward, exact=True)
    assert k.shape == ke.shape
    # HEALPix n != 0: the m-recursion's NaN-dropped rings (cos theta = 1/3 at nside 2)
    if sampling == "healpix":
        ok = ~np.isclose(np.abs(np.cos(geo.thetas(L, "healpix", nside))), 1 / 3)
        k, ke = k[:, ok], ke[:, ok]
    assert np.max(np.abs(k - ke)) < 1e-12 * np.max(np.abs(ke))


def test_torch_twin_matches_the_oracle_recursion():
    """The CPU baseline's torch twin computes what the oracle's latitudinal
    step computes (the first m-steps of the positive half)."""
    torch = pytest.importorskip("torch")
    from oracle import torch_twin

    nside, L = 16, 32
    beta = geo.thetas(L, "healpix", nside)
    rng = np.random.default_rng(3)
    ftm = rng.standard_normal((len(beta), 2 * L)) + 1j * rng.standard_normal((len(beta), 2 * L))
    pre = torch_twin.precompute(beta, L)
    ref_pre = pm.precomputes(L, 0, "healpix", nside, True, 0, beta)
    assert np.max(np.abs(pre["lren"].numpy() - ref_pre[0][1])) < 1e-12 * np.max(np.abs(ref_pre[0][1]))
    got = torch_twin.forward_latitudinal_sample(ftm, pre, L, L).numpy()
    flm = pm.forward_latitudinal_step(ftm, beta, L, 0, nside, "healpix", True)
    ref = np.stack([flm[:, L - 1 + m] for m in range(L - 1, -1, -1)])
    assert max_rel(got, ref) < 1e-13
