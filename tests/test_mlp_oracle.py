"""Pins of the learned preconditioner's CPU oracle (no GPU).

* BF16 rounding against torch's float32 -> bfloat16 conversion (round to
  nearest even), including ties and extremes.
* The hand-derived gradient against central finite differences of the loss
  (roundings switched off, float64), so a dropped term or a transposed weight
  fails.
* FP16 rounding against numpy's float16 conversion and torch's float16.
* ReLU and its derivative against finite differences.
* Feature rows: RMS scaling and (l, m) enumeration against a direct loop.
"""

import numpy as np
import pytest
import torch

from oracle import mlp


def test_bf16_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 10000),
                        np.array([1.0, 1.00390625, 1.01171875, -2.5, 3e38, 1e-38, 0.0], np.float32)])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    got = mlp.bf16(x)
    fin = np.isfinite(ref)
    assert np.array_equal(got[fin], ref[fin])


def test_f16_matches_torch():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-6, 5, 10000),
                        [1.0, 1.0 + 2 ** -11, 1.0 + 3 * 2 ** -11, -2.5, 6e4, 0.0]])
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.float16).to(torch.float64).numpy()
    got = mlp.f16(x.astype(np.float32))
    assert np.array_equal(got, ref)


def test_relu_and_grad():
    x = np.linspace(-6, 6, 241)
    assert np.array_equal(mlp.relu(x), np.where(x > 0, x, 0.0))
    xs = x[np.abs(x) > 1e-3]
    h = 1e-6
    fd = (mlp.relu(xs + h) - mlp.relu(xs - h)) / (2 * h)
    assert np.allclose(mlp.relu_grad(xs), fd, atol=1e-8)


def test_grad_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = mlp.init_params(rng).astype(np.float64)
    p[mlp.H * 4:mlp.H * 5] = rng.standard_normal(mlp.H) * 0.1        # b1
    p[-2:] = [0.3, -0.2]                                              # b3
    X = rng.standard_normal((7, 4))
    Y = rng.standard_normal((7, 2))
    w = rng.uniform(0.5, 1.5, 7)
    g = mlp.grad(p, X, Y, w, loss_scale=0.7, quantize=False)
    seg = mlp.segments()
    picks = {"W1": 6, "b1": 3, "W2": 8, "b2": 3, "W3": 4, "b3": 2}
    idx = np.concatenate([lo + rng.choice(hi - lo, picks[k], replace=False)
                          for k, (lo, hi) in seg.items()])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = 0.7 * (mlp.loss(pp, X, Y, w, quantize=False) - mlp.loss(pm, X, Y, w, quantize=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


@pytest.mark.parametrize("reality", [True, False])
def test_features_direct_loop(reality):
    rng = np.random.default_rng(2)
    L, lmin = 6, 1
    h = rng.standard_normal((L, 2 * L - 1)) + 1j * rng.standard_normal((L, 2 * L - 1))
    e = rng.standard_normal((L, 2 * L - 1)) + 1j * rng.standard_normal((L, 2 * L - 1))
    X, Y, s = mlp.features(h, e, L, reality, lmin)
    vals = []
    for l in range(lmin, L):
        for m in range(0 if reality else -l, l + 1):
            vals.append(h[l, L - 1 + m])
    assert np.isclose(s, np.sqrt(np.mean(np.abs(vals) ** 2)))
    # row of (l, m) = (3, 2)
    i = (3 * 4 // 2 + 2) if reality else (9 + 3 + 2)
    assert np.allclose(X[i], [h[3, L + 1].real / s, h[3, L + 1].imag / s, 3 / L, 2 / L])
    assert np.allclose(Y[i], [e[3, L + 1].real / s, e[3, L + 1].imag / s])
    assert np.all(X[0] == 0)     # l = 0 < lmin


def test_precondition_reality_symmetry():
    rng = np.random.default_rng(3)
    L = 5
    p = mlp.init_params(rng)
    h = rng.standard_normal((L, 2 * L - 1)) + 1j * rng.standard_normal((L, 2 * L - 1))
    out = mlp.precondition(p, h, L, True)
    for l in range(L):
        assert out[l, L - 1].imag == 0
        for m in range(1, l + 1):
            assert np.isclose(out[l, L - 1 - m], (-1) ** m * np.conj(out[l, L - 1 + m]))
