"""A torch-CPU twin of the oracle's latitudinal m-recursion (CPU baseline).

Test infrastructure only -- see ``oracle/__init__.py``.  Used by the
``cpu_baseline`` leg of ``bench.py`` as the stand-in for the reference's JAX
CPU path (jax is not installable here): the same array program as
s2fft/transforms/otf_recursions.py:591-886 (the JAX forward latitudinal step,
vectorised over (theta, l), one m per step, log-space renormalisation after
every step) in float64 torch tensors, so the CPU runs it on all host threads.
``forward_latitudinal_sample`` runs the first ``msteps`` m-steps of the
positive-m half on the given rings and returns the projected coefficients of
those m; every m-step costs the same, so a sample scales linearly.
"""

import numpy as np
import torch

from . import pm


def precompute(beta, L, spin=0):
    """The recursion's precomputes for the rings beta (price_mcewen.py:118-255,
    as pm.precomputes) in torch: the streaming log-ratio sums of the boundary
    values (212-243) in closed form, lrenorm[t, l] = 2 l log|cos(beta_t/2)| +
    sum_{j=2}^{h_l} (ratio_j(l) + log|tan(beta_t/2)|), h_l = l - mm + 1 (the
    positive-m half), as a cumulative sum; cpi, cp2 and vsign as in pm."""
    beta = np.asarray(beta, dtype=np.float64)
    mm = -spin
    tb = torch.as_tensor(beta, dtype=torch.float64)
    el = torch.arange(L, dtype=torch.float64)
    lt = torch.log(torch.abs(torch.tan(-tb / 2.0)))
    lc2 = torch.log(torch.abs(torch.cos(tb / 2.0)))
    i_all = torch.arange(2 * L + 1, dtype=torch.float64)[:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = torch.log((2 * el[None, :] + 2 - i_all) / (i_all - 1)) / 2
    ratio[:2] = 0.0
    S = torch.cumsum(ratio, dim=0)                  # S[h] = sum_{j=2}^{h} ratio[j]
    h = (el - mm + 1).long()
    valid = (h >= 1) & (h <= L + abs(mm) + 1)
    hc = h.clamp(1, 2 * L)
    lren = (2 * el[None, :] * lc2[:, None] + S[hc, torch.arange(L)][None, :]
            + (hc - 1).to(torch.float64)[None, :] * lt[:, None])
    lren = torch.where(valid[None, :], lren, torch.zeros_like(lren))
    lren = torch.nan_to_num(lren, nan=0.0, posinf=0.0, neginf=0.0)
    # cpi, cp2 (181-201) and vsign (207-210): O(L^2), as pm.precomputes
    eln = np.arange(L)
    with np.errstate(divide="ignore", invalid="ignore"):
        cpi = np.zeros((L + 1, L))
        cp2 = np.zeros((L + 1, L))
        cpi[0] = 2.0 / np.sqrt(2 * eln)
        for m in range(2, L + 1):
            cpi[m - 1] = 2.0 / np.sqrt(m * (2 * eln + 1 - m))
            cp2[m - 1] = cpi[m - 1] / cpi[m - 2]
        for k in range(L):
            cpi[:, k] = np.roll(cpi[:, k], L - k - 1)
            cp2[:, k] = np.roll(cp2[:, k], L - k - 1)
    msign = np.concatenate(((-1.0) ** np.arange(L - 1), np.ones(L)))
    vsign = np.outer(msign, (-1.0) ** np.abs(mm + eln))
    vsign[: L - 1] *= (-1.0) ** abs(mm + 1 + L)
    t = lambda a: torch.as_tensor(np.nan_to_num(a, nan=0.0, posinf=0.0, neginf=0.0),
                                  dtype=torch.float64)
    return {"lren": lren, "vsign": t(vsign), "cpi": t(cpi), "cp2": t(cp2),
            "idx": torch.arange(L)[None, :].expand(len(beta), L), "c": t(np.cos(beta))[:, None],
            "s": t(np.sin(beta))[:, None], "spin": spin}


def forward_latitudinal_sample(ftm, pre, L, msteps):
    """flm[l, m] = sum_t ftm[t, m] d^l_{m,-spin}(beta_t) for m = L-1, L-2, ...
    (msteps of them), as (msteps, L) complex128; ftm: (nt, 2L) complex128 in
    the reference ftm layout (column L + m); pre: ``precompute(beta, L, spin)``."""
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64)
    lren, vsign, cpi, cp2, idx = pre["lren"].clone(), pre["vsign"], pre["cpi"], pre["cp2"], pre["idx"]
    c, s, spin = pre["c"], pre["s"], pre["spin"]
    nt = c.shape[0]
    el = torch.arange(L, dtype=torch.float64)
    omc = 1.0 - c
    half = el - (-spin) + 1
    fr = t(ftm.real)
    fi = t(ftm.imag)
    out = torch.zeros((msteps, L), dtype=torch.complex128)
    sind, sgn = 2 * L - 2, -1
    lind = L - 1
    it0 = torch.ones((nt, L), dtype=torch.float64)
    it1 = torch.ones((nt, L), dtype=torch.float64)
    lamb = (omc * (el + 1) + c * (2 - L + el) - half) / s
    it1[:, lind:] = cpi[0, lind:] * it0[:, lind:] * lamb[:, lind:]
    entry = torch.zeros((nt, L), dtype=torch.float64)
    for k in range(msteps):
        m = L - 1 - k
        if k == 0:  # |m| = L-1: only l = L-1
            D = torch.zeros((nt, L), dtype=torch.float64)
            D[:, lind:] = it0[:, lind:] * vsign[sind, lind:] * torch.exp(lren[:, lind:])
        elif k == 1:  # |m| = L-2: l in {L-2, L-1}
            D = torch.zeros((nt, L), dtype=torch.float64)
            D[:, lind - 1:] = (it1[:, lind - 1:] * vsign[sind + sgn, lind - 1:]
                               * torch.exp(lren[:, lind - 1:]))
        else:
            active = idx >= L - k - 1
            lamb = (omc * (el + 1) + c * (k - L + el + 1) - half) / s
            entry = torch.where(active, cpi[k - 1] * it1 * lamb - cp2[k - 1] * it0, entry)
            entry[:, L - 1 - k] = 1.0
            D = entry * vsign[sind + sgn * k] * torch.exp(lren)
            mag = entry.abs()
            it0 = torch.where(active, it1 / mag, it0)
            it1 = torch.where(active, entry / mag, it1)
            lren = torch.where(active, lren + torch.log(mag), lren)
        D = torch.nan_to_num(D, nan=0.0, posinf=0.0, neginf=0.0)
        out[k] = torch.complex(fr[:, L + m] @ D, fi[:, L + m] @ D)
    return out
