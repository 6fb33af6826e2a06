"""Learned preconditioner for the HEALPix iterative forward transform.

A per-(l, m) MLP maps the harmonic residual h = F(f - I(flm)) (scaled by its
RMS s) and the position (l/L, m/L) to a correction of flm:

    delta[l, m] = s * MLP(h[l, m] / s, l / L, m / L)

Two hidden layers of 64 with ReLU; FP16 weights, features and activations,
FP32 accumulation (s2h_mlp_* in the C ABI).  Training data come from the plain
FP64 transform: seeded band-limited coefficients flm*, maps f = I(flm*),
Jacobi iterates x_k, residuals h_k = F(f - I(x_k)) and errors e_k = flm* - x_k;
the network learns e_k from h_k.  Training runs on the GPU with the
library's forward/backward kernels and Adam.

The reference has no counterpart (its refinement is plain Jacobi,
s2fft/utils/iterative_refinement.py:9-46); Jacobi stays the fallback and
judge: a learned step is kept for a map only where it lowers that map's
FP64-held residual norm, else the map takes the Jacobi step
(``refine._judged_step``).
"""

import math
from pathlib import Path

import numpy as np
import torch

from .. import _lib
from ..transforms import spherical as sph

WEIGHTS = Path(__file__).resolve().parent / "healpix_spin0.npy"


def param_count():
    return int(_lib.lib().s2h_mlp_param_count())


def hidden_width():
    """H of the library's network: param_count() = H^2 + 8 H + 2."""
    return int(round(math.sqrt(param_count() + 14))) - 4


def init_params(seed=0, device="cuda"):
    """Kaiming-uniform-like initialisation (biases zero, last layer small)."""
    H = hidden_width()
    g = np.random.default_rng(seed)
    W1 = g.uniform(-1, 1, (H, 4)) / 2.0
    W2 = g.uniform(-1, 1, (H, H)) * math.sqrt(3.0 / H)
    W3 = g.uniform(-1, 1, (2, H)) * math.sqrt(3.0 / H) * 0.1
    p = np.concatenate([W1.ravel(), np.zeros(H), W2.ravel(), np.zeros(H), W3.ravel(),
                        np.zeros(2)]).astype(np.float32)
    assert p.size == param_count()
    return torch.from_numpy(p).to(device)


def _stream(t):
    return torch.cuda.current_stream(t.device).cuda_stream


def _lmin(spin, L_lower):
    return max(abs(spin), L_lower)


class MLPPreconditioner:
    """Holds the FP32 master parameters on the device and applies the network."""

    def __init__(self, params):
        self.params = params.to(torch.float32).contiguous()
        assert self.params.numel() == param_count()

    @classmethod
    def load(cls, path=WEIGHTS, device="cuda"):
        return cls(torch.from_numpy(np.load(path, allow_pickle=False)).to(device))

    def save(self, path=WEIGHTS):
        np.save(path, self.params.cpu().numpy())

    def apply(self, h, L, reality, lmin=0, packed=False):
        """delta for a residual h of the same shape: (B, L, 2L-1) complex128, or
        (B, L(L+1)/2) in the packed m-major layout (``packed``, real maps)."""
        h = h.contiguous()
        B = h.shape[0]
        out = torch.empty_like(h)
        work = _mlp_work(B, h.device)
        mode = 2 if packed else int(bool(reality))
        _lib.check(_lib.lib().s2h_mlp_precondition(
            self.params.data_ptr(), h.data_ptr(), out.data_ptr(), L, mode, lmin,
            B, work.data_ptr(), _stream(h)), "mlp precondition")
        return out

    def apply_dot(self, h, L, reality, lmin=0, packed=False):
        """(delta, <delta, h> per map) in one pass: the dot is the one
        ``flm_dots_hip(delta, [h], ...)`` gives (c = 2 for the m > 0 rows of
        real maps), summed in the network's own fixed order."""
        h = h.contiguous()
        B = h.shape[0]
        out = torch.empty_like(h)
        dot = torch.empty(max(B, 1), dtype=torch.float64, device=h.device)
        work = _mlp_work(B, h.device)
        mode = 2 if packed else int(bool(reality))
        _lib.check(_lib.lib().s2h_mlp_precondition_dot(
            self.params.data_ptr(), h.data_ptr(), out.data_ptr(), L, mode, lmin,
            B, dot.data_ptr(), work.data_ptr(), _stream(h)), "mlp precondition")
        return out, dot[:B]

    def __call__(self, h, r=None, flm=None, L=None, reality=True, lmin=0):
        return self.apply(h, L if L is not None else h.shape[1], reality, lmin)


def _mlp_work(B, device):
    """Workspace of the RMS pass (fixed-order partial sums per map)."""
    n = _lib.lib().s2h_mlp_workspace_bytes(max(B, 1)) // 8
    return torch.empty(n, dtype=torch.float64, device=device)


def features(h, e, L, reality, lmin=0):
    """Training rows X (B*P, 4), Y (B*P, 2) from residual h and error e."""
    h, e = h.contiguous(), e.contiguous()
    B = h.shape[0]
    P = L * (L + 1) // 2 if reality else L * L
    X = torch.empty((B * P, 4), dtype=torch.float32, device=h.device)
    Y = torch.empty((B * P, 2), dtype=torch.float32, device=h.device)
    work = _mlp_work(B, h.device)
    _lib.check(_lib.lib().s2h_mlp_features(h.data_ptr(), e.data_ptr(), L, int(bool(reality)),
                                           lmin, B, X.data_ptr(), Y.data_ptr(),
                                           work.data_ptr(), _stream(h)), "mlp features")
    return X, Y


class _GradWork:
    def __init__(self, nwg, device):
        self.nwg = nwg
        nbytes = int(_lib.lib().s2h_mlp_grad_workspace_bytes(nwg))
        self.buf = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.grad = torch.empty(param_count(), dtype=torch.float32, device=device)
        self.loss = torch.zeros(1, dtype=torch.float64, device=device)


def grad(params, X, Y, row_weight=None, loss_scale=1.0, nwg=256, work=None):
    """(grad, loss) of loss_scale * sum w |MLP(X) - Y|^2 (loss unscaled)."""
    work = work or _GradWork(nwg, params.device)
    n = X.shape[0]
    _lib.check(_lib.lib().s2h_mlp_grad(
        params.data_ptr(), X.data_ptr(), Y.data_ptr(),
        row_weight.data_ptr() if row_weight is not None else None, n, float(loss_scale),
        work.grad.data_ptr(), work.loss.data_ptr(), work.buf.data_ptr(), work.nwg,
        _stream(params)), "mlp grad")
    return work.grad, work.loss


class Adam:
    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8):
        self.params = params
        self.m1 = torch.zeros_like(params)
        self.m2 = torch.zeros_like(params)
        self.lr, self.betas, self.eps = lr, betas, eps
        self.step_count = 0

    def step(self, g, lr=None):
        self.step_count += 1
        _lib.check(_lib.lib().s2h_mlp_adam(
            self.params.data_ptr(), g.data_ptr(), self.m1.data_ptr(), self.m2.data_ptr(),
            self.step_count, float(lr if lr is not None else self.lr), float(self.betas[0]),
            float(self.betas[1]), float(self.eps), _stream(self.params)), "mlp adam")


def make_training_rows(nside, n_maps, seed=1000, batch=16, iters=(0, 1, 2), keep=1.0,
                       log=None):
    """Rows (X, Y) from Jacobi iterates of the plain FP64 HEALPix transform.

    Maps: seeded band-limited real spin-0 signals at L = 2 nside (flm entries
    complex N(0,1) for m > 0, real N(0,1) for m = 0).  ``keep`` < 1 keeps a
    seeded random fraction of the rows of every (map batch, iterate).
    """
    L = 2 * nside
    dev = torch.device("cuda")
    plan = sph.get_plan("healpix", L, nside, 0, 0, True, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    Xs, Ys = [], []
    ell = torch.arange(L, device=dev)[:, None]
    mm = torch.arange(-(L - 1), L, device=dev)[None, :]
    pos = (mm >= 0) & (mm <= ell)
    for b0 in range(0, n_maps, batch):
        B = min(batch, n_maps - b0)
        re = torch.randn((B, L, 2 * L - 1), generator=gen, device=dev, dtype=torch.float64)
        im = torch.randn((B, L, 2 * L - 1), generator=gen, device=dev, dtype=torch.float64)
        flm = torch.complex(re, torch.where(mm == 0, torch.zeros_like(im), im))
        flm = torch.where(pos, flm, torch.zeros_like(flm))
        # conjugate symmetry for m < 0
        sign = torch.where(mm % 2 == 0, 1.0, -1.0).to(torch.float64)
        mirror = torch.flip(flm, dims=[2]).conj() * sign
        flm = torch.where(mm < 0, mirror, flm)
        f = sph.inverse_hip(flm, L, 0, nside, "healpix", True, plan=plan).real.contiguous()
        x = sph.forward_hip(f, L, 0, nside, "healpix", True, plan=plan)
        for k in range(max(iters) + 1):
            r = f - sph.inverse_hip(x, L, 0, nside, "healpix", True, plan=plan).real
            h = sph.forward_hip(r.contiguous(), L, 0, nside, "healpix", True, plan=plan)
            if k in iters:
                X, Y = features(h, flm - x, L, True)
                if keep < 1.0:
                    sel = torch.rand(X.shape[0], generator=gen, device=dev) < keep
                    X, Y = X[sel], Y[sel]
                Xs.append(X)
                Ys.append(Y)
            x = x + h
        if log and (b0 // batch) % 16 == 0:
            log(f"[mlp] training data: {b0 + B}/{n_maps} maps")
    return torch.cat(Xs), torch.cat(Ys)


def train(X, Y, steps=2000, rows_per_step=65536, lr=2e-3, seed=0, log=None, params=None):
    """Adam on mean squared error over random minibatches; returns the model."""
    dev = X.device
    params = init_params(seed, dev) if params is None else params.clone()
    opt = Adam(params, lr=lr)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed + 17)
    work = _GradWork(256, dev)
    n = X.shape[0]
    hist = []
    for it in range(steps):
        idx = torch.randint(0, n, (rows_per_step,), generator=gen, device=dev)
        xb, yb = X[idx].contiguous(), Y[idx].contiguous()
        g, loss = grad(params, xb, yb, None, 1.0 / (2 * rows_per_step), work=work)
        # cosine decay
        cur = lr * 0.5 * (1 + math.cos(math.pi * it / steps))
        opt.step(g, cur)
        if it % 100 == 0 or it == steps - 1:
            l = float(loss.item()) / (2 * rows_per_step)
            hist.append((it, l))
            if log:
                log(f"[mlp] step {it}: loss {l:.4e}")
    return MLPPreconditioner(params), hist
