"""CPU restatement of the learned preconditioner's network (test oracle).

Test infrastructure only -- see ``oracle/__init__.py``.

No reference counterpart: the reference refines with plain Jacobi steps
(s2fft/utils/iterative_refinement.py:9-46).  The network is
    z1 = W1h h(x) + b1,  a1 = h(relu(z1)),  z2 = W2h a1 + b2,
    a2 = h(relu(z2)),    y  = W3h a2 + b3,   Wh = h(W), h = round to fp16
evaluated here in float64 (the FP16 roundings applied exactly), and its
gradient by hand-derived backpropagation with the roundings treated as the
identity.  The gradient is pinned in the tests by central finite
differences of ``loss`` (with the roundings switched off), and the feature
construction by a direct loop.
"""

import numpy as np

H = 64
N_PARAMS = H * 4 + H + H * H + H + 2 * H + 2


def segments():
    """(start, end) of W1, b1, W2, b2, W3, b3 in the packed parameters."""
    sizes = [("W1", 4 * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", 2 * H), ("b3", 2)]
    out, o = {}, 0
    for name, n in sizes:
        out[name] = (o, o + n)
        o += n
    return out

def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even)."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out).astype(np.float32)


def unpack(p):
    p = np.asarray(p)
    o = 0
    W1 = p[o:o + 4 * H].reshape(H, 4); o += 4 * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + 2 * H].reshape(2, H); o += 2 * H
    b3 = p[o:o + 2]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def init_params(rng):
    """Kaiming-uniform-like init (biases zero, last layer small)."""
    W1 = rng.uniform(-1, 1, (H, 4)) / np.sqrt(4)
    W2 = rng.uniform(-1, 1, (H, H)) * np.sqrt(3.0 / H)
    W3 = rng.uniform(-1, 1, (2, H)) * np.sqrt(3.0 / H) * 0.1
    return pack(W1, np.zeros(H), W2, np.zeros(H), W3, np.zeros(2)).astype(np.float32)


def f16(x):
    """Round to the nearest IEEE half (ties to even), as float64."""
    return np.asarray(x, dtype=np.float64).astype(np.float16).astype(np.float64)


def relu(x):
    return np.maximum(np.asarray(x, dtype=np.float64), 0.0)


def relu_grad(x):
    return (np.asarray(x, dtype=np.float64) > 0).astype(np.float64)


def forward(p, X, quantize=True):
    """Returns y (n, 2) and the intermediates."""
    W1, b1, W2, b2, W3, b3 = (np.asarray(t, dtype=np.float64) for t in unpack(p))
    qh = f16 if quantize else (lambda t: t)
    W1q, W2q, W3q = qh(W1), qh(W2), qh(W3)
    X = qh(np.asarray(X, dtype=np.float64))
    z1 = X @ W1q.T + b1
    a1 = qh(relu(z1))
    z2 = a1 @ W2q.T + b2
    a2 = qh(relu(z2))
    y = a2 @ W3q.T + b3
    return y, (X, z1, a1, z2, a2, W1q, W2q, W3q)


def loss(p, X, Y, w=None, quantize=True):
    y, _ = forward(p, X, quantize)
    d = (y - np.asarray(Y, dtype=np.float64)) ** 2
    if w is not None:
        d = d * np.asarray(w, dtype=np.float64)[:, None]
    return float(d.sum())


def grad(p, X, Y, w=None, loss_scale=1.0, quantize=True):
    """Gradient of loss_scale * sum_r w_r |y_r - Y_r|^2, packed like p."""
    y, (X, z1, a1, z2, a2, W1q, W2q, W3q) = forward(p, X, quantize)
    wt = np.ones(len(X)) if w is None else np.asarray(w, dtype=np.float64)
    dy = 2.0 * (y - np.asarray(Y, dtype=np.float64)) * wt[:, None] * loss_scale
    gW3 = dy.T @ a2
    gb3 = dy.sum(0)
    dz2 = (dy @ W3q) * relu_grad(z2)
    gW2 = dz2.T @ a1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2q) * relu_grad(z1)
    gW1 = dz1.T @ X
    gb1 = dz1.sum(0)
    return pack(gW1, gb1, gW2, gb2, gW3, gb3)


def features(h, e, L, reality, lmin=0):
    """Training rows (X, Y) from residual h and error e (each (L, 2L-1))."""
    rows = [(l, m) for l in range(L) for m in (range(0, l + 1) if reality else range(-l, l + 1))]
    valid = [(l, m) for (l, m) in rows if l >= lmin]
    s = np.sqrt(np.mean([abs(h[l, L - 1 + m]) ** 2 for (l, m) in valid])) if valid else 1.0
    if s == 0:
        s = 1.0
    X = np.zeros((len(rows), 4))
    Y = np.zeros((len(rows), 2))
    for i, (l, m) in enumerate(rows):
        if l < lmin:
            continue
        v, t = h[l, L - 1 + m], e[l, L - 1 + m]
        X[i] = (v.real / s, v.imag / s, l / L, m / L)
        Y[i] = (t.real / s, t.imag / s)
    return X, Y, s


def precondition(p, h, L, reality, lmin=0):
    """Correction s * y(X) written into an (L, 2L-1) array (see the C ABI)."""
    X, _, s = features(h, h, L, reality, lmin)
    y, _ = forward(p, X)
    out = np.zeros((L, 2 * L - 1), dtype=np.complex128)
    rows = [(l, m) for l in range(L) for m in (range(0, l + 1) if reality else range(-l, l + 1))]
    for i, (l, m) in enumerate(rows):
        if l < lmin:
            continue
        re, im = s * y[i, 0], s * y[i, 1]
        if reality and m == 0:
            im = 0.0
        out[l, L - 1 + m] = re + 1j * im
        if reality and m > 0:
            out[l, L - 1 - m] = (-1) ** m * (re - 1j * im)
    return out
