"""Learned preconditioner kernels against the CPU oracle (needs a GPU).

Tolerances: every layer takes FP16 operands in both the kernel and the
oracle (the same roundings), but the kernel accumulates in FP32 (the oracle
in float64) and rounds the FP32 pre-activations to FP16 where the oracle
rounds float64 ones, so an occasional activation differs by one FP16 unit:
the tests allow 2e-3 of the output scale.  Gradients: FP32 accumulation over the rows vs FP64,
1e-3 of each block's largest entry.
"""

import numpy as np
import pytest
import torch

from oracle import mlp as omlp

from s2fft_amd import _lib
from s2fft_amd.learned import precond as lp

pytestmark = pytest.mark.gpu


def _stream():
    return torch.cuda.current_stream().cuda_stream


def test_mfma_bf16_layout():
    """V_MFMA_F32_16X16X32_BF16: A[i][k] at lane i + 16 (k // 8), element k % 8;
    B[k][j] at lane j + 16 (k // 8), element k % 8; register r of lane l holds
    C[4 (l // 16) + r][l % 16]."""
    rng = np.random.default_rng(7)
    A = omlp.bf16(rng.standard_normal((16, 32)).astype(np.float32))
    B = omlp.bf16(rng.standard_normal((32, 16)).astype(np.float32))
    lanes = np.arange(64)
    a = np.stack([A[l % 16, 8 * (l // 16):8 * (l // 16) + 8] for l in lanes])
    b = np.stack([B[8 * (l // 16):8 * (l // 16) + 8, l % 16] for l in lanes])
    ab = (a.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    bb = (b.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    da = torch.from_numpy(ab.astype(np.int16)).cuda()
    db = torch.from_numpy(bb.astype(np.int16)).cuda()
    out = torch.zeros(256, dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().s2h_probe_mfma_bf16(da.data_ptr(), db.data_ptr(), out.data_ptr(),
                                              _stream()))
    got = out.cpu().numpy().reshape(64, 4)
    C = A.astype(np.float64) @ B.astype(np.float64)
    expect = np.array([[C[4 * (l // 16) + r, l % 16] for r in range(4)] for l in range(64)])
    assert np.allclose(got, expect, atol=1e-4)


def _params(seed):
    rng = np.random.default_rng(seed)
    p = omlp.init_params(rng)
    p[omlp.H * 4:omlp.H * 5] = rng.standard_normal(omlp.H) * 0.1
    lo, hi = omlp.segments()["b2"]
    p[lo:hi] = rng.standard_normal(omlp.H) * 0.1
    p[-2:] = [0.05, -0.03]
    return p.astype(np.float32)


@pytest.mark.parametrize("L,reality,lmin,B", [(12, True, 0, 2), (9, False, 2, 3), (33, True, 1, 1),
                                               (400, True, 3, 1)])
def test_precondition_matches_oracle(L, reality, lmin, B):
    rng = np.random.default_rng(L)
    p = _params(L)
    h = rng.standard_normal((B, L, 2 * L - 1)) + 1j * rng.standard_normal((B, L, 2 * L - 1))
    model = lp.MLPPreconditioner(torch.from_numpy(p).cuda())
    got = model.apply(torch.from_numpy(h).cuda(), L, reality, lmin).cpu().numpy()
    for b in range(B):
        ref = omlp.precondition(p, h[b], L, reality, lmin)
        scale = np.abs(ref).max()
        assert np.abs(got[b] - ref).max() <= 2e-3 * scale
        if reality:
            assert np.all(got[b][:, L - 1].imag == 0)


@pytest.mark.parametrize("L,reality,packed,lmin,B", [(512, True, True, 0, 4), (40, True, False, 2, 3),
                                                     (33, False, False, 1, 2)])
def test_precondition_dot_matches_flm_dots(L, reality, packed, lmin, B):
    """s2h_mlp_precondition_dot: the same delta bit for bit, and <delta, h>
    equal to s2h_flm_dots' to 1e-12 relative (another fixed summation order)."""
    from s2fft_amd.transforms import spherical as sph
    g = torch.Generator(device="cuda")
    g.manual_seed(L)
    shape = (B, L * (L + 1) // 2) if packed else (B, L, 2 * L - 1)
    h = torch.randn(shape, generator=g, dtype=torch.complex128, device="cuda")
    if not packed:
        h = torch.from_numpy(np.stack([omlp_full(hb.cpu().numpy(), L, reality)
                                       for hb in h])).cuda()
    model = lp.MLPPreconditioner.load(device=torch.device("cuda"))
    ref = model.apply(h, L, reality, lmin, packed=packed)
    got, dot = model.apply_dot(h, L, reality, lmin, packed=packed)
    assert torch.equal(got, ref)
    want = sph.flm_dots_hip(ref, [h], L if packed else 0)[0]
    assert torch.allclose(dot, want, rtol=1e-12, atol=0)
    assert torch.all(dot.abs() > 0)


def omlp_full(h, L, reality):
    """A full-layout input (L, 2L-1) that holds only valid rows: |m| <= l, and
    for real maps the conjugate-symmetric m < 0 half."""
    ell = np.arange(L)[:, None]
    mm = np.arange(-(L - 1), L)[None, :]
    h = np.where(np.abs(mm) <= ell, h, 0)
    if reality:
        sign = np.where(mm % 2 == 0, 1.0, -1.0)
        h = np.where(mm == 0, h.real + 0j, h)
        h = np.where(mm < 0, np.flip(h, axis=1).conj() * sign, h)
    return h


def test_precondition_deterministic():
    """The residual's RMS is summed in a fixed order: equal outputs, bit for
    bit, on every call (L = 512 packed: more rows than the kernels' grids)."""
    L, B = 512, 4
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    h = torch.randn((B, L * (L + 1) // 2), generator=g, dtype=torch.complex128, device="cuda")
    model = lp.MLPPreconditioner.load(device=torch.device("cuda"))
    first = model.apply(h, L, True, 0, packed=True).clone()
    for _ in range(3):
        assert torch.equal(model.apply(h, L, True, 0, packed=True), first)


@pytest.mark.parametrize("reality", [True, False])
def test_features_match_oracle(reality):
    rng = np.random.default_rng(11)
    L, lmin, B = 10, 2, 2
    h = rng.standard_normal((B, L, 2 * L - 1)) + 1j * rng.standard_normal((B, L, 2 * L - 1))
    e = rng.standard_normal((B, L, 2 * L - 1)) + 1j * rng.standard_normal((B, L, 2 * L - 1))
    X, Y = lp.features(torch.from_numpy(h).cuda(), torch.from_numpy(e).cuda(), L, reality, lmin)
    X, Y = X.cpu().numpy(), Y.cpu().numpy()
    P = L * (L + 1) // 2 if reality else L * L
    for b in range(B):
        Xr, Yr, _ = omlp.features(h[b], e[b], L, reality, lmin)
        assert np.allclose(X[b * P:(b + 1) * P], Xr, rtol=1e-6, atol=1e-6)
        assert np.allclose(Y[b * P:(b + 1) * P], Yr, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("n,nwg", [(1000, 7), (77, 256), (4096, 64)])
def test_grad_matches_oracle(n, nwg):
    rng = np.random.default_rng(n)
    p = _params(n)
    X = rng.standard_normal((n, 4)).astype(np.float32)
    Y = rng.standard_normal((n, 2)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    dp = torch.from_numpy(p).cuda()
    g, loss = lp.grad(dp, torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda(),
                      torch.from_numpy(w).cuda(), 1.0 / n, nwg=nwg)
    g = g.cpu().numpy().astype(np.float64)
    ref = omlp.grad(p, X, Y, w, 1.0 / n)
    for lo, hi in omlp.segments().values():
        s = np.abs(ref[lo:hi]).max()
        assert np.abs(g[lo:hi] - ref[lo:hi]).max() <= 1e-3 * s, (lo, hi)
    assert abs(float(loss.item()) - omlp.loss(p, X, Y, w)) <= 1e-4 * omlp.loss(p, X, Y, w)


def test_grad_deterministic():
    rng = np.random.default_rng(3)
    p = torch.from_numpy(_params(3)).cuda()
    X = torch.from_numpy(rng.standard_normal((5000, 4)).astype(np.float32)).cuda()
    Y = torch.from_numpy(rng.standard_normal((5000, 2)).astype(np.float32)).cuda()
    g1 = lp.grad(p, X, Y, None, 1.0, nwg=32)[0].clone()
    g2 = lp.grad(p, X, Y, None, 1.0, nwg=32)[0].clone()
    assert torch.equal(g1, g2)


def test_adam_matches_numpy():
    rng = np.random.default_rng(4)
    n = lp.param_count()
    p = rng.standard_normal(n).astype(np.float32)
    dp = torch.from_numpy(p.copy()).cuda()
    opt = lp.Adam(dp, lr=1e-2)
    m1 = np.zeros(n, np.float64)
    m2 = np.zeros(n, np.float64)
    pr = p.astype(np.float64)
    for t in range(1, 4):
        g = rng.standard_normal(n).astype(np.float32)
        opt.step(torch.from_numpy(g).cuda())
        m1 = 0.9 * m1 + 0.1 * g
        m2 = 0.999 * m2 + 0.001 * g.astype(np.float64) ** 2
        pr = pr - 1e-2 * (m1 / (1 - 0.9 ** t)) / (np.sqrt(m2 / (1 - 0.999 ** t)) + 1e-8)
    assert np.allclose(dp.cpu().numpy(), pr, rtol=1e-5, atol=1e-6)


def test_training_fits_a_linear_map():
    """The training loop drives the loss down on a target the network can fit."""
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    X = torch.randn((50000, 4), generator=g, device="cuda")
    X[:, 2:] = torch.rand((50000, 2), generator=g, device="cuda")
    scale = 1.0 + 0.1 * X[:, 2:3]
    Y = (X[:, :2] * scale).contiguous()
    model, hist = lp.train(X.contiguous(), Y, steps=300, rows_per_step=8192, lr=3e-3)
    assert hist[-1][1] < 0.05 * hist[0][1]
