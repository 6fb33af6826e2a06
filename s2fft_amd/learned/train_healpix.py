"""Train the learned preconditioner for HEALPix spin-0 forward refinement.

    python -m s2fft_amd.learned.train_healpix [--nside 256] [--maps 4096]
        [--steps 2000] [--out s2fft_amd/learned/healpix_spin0.npy]

Training data come from the plain FP64 transform on the GPU (seeded
band-limited maps at nside, L = 2 nside; residuals and errors of the first
Jacobi iterates), and the network is trained on the GPU by the library's own
forward / backward kernels with Adam.  The script then compares, on held-out
maps at ``--eval-nside``, the steps the mixed refinement needs to reach 1e-7
with and without the network, and prints one JSON line.
"""

import argparse
import json
import sys
import time

import torch

from . import precond as lp
from ..transforms import refine
from ..transforms import spherical as sph


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def eval_gain(model, nside, n_maps=8, seed=9000, tol=1e-7, max_iter=12, precision="bf16",
              refresh_every=3):
    """Steps to tol of the mixed refinement with and without the network."""
    L = 2 * nside
    dev = torch.device("cuda")
    plan = sph.get_plan("healpix", L, nside, 0, 0, True, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    ell = torch.arange(L, device=dev)[:, None]
    mm = torch.arange(-(L - 1), L, device=dev)[None, :]
    re = torch.randn((n_maps, L, 2 * L - 1), generator=gen, device=dev, dtype=torch.float64)
    im = torch.randn((n_maps, L, 2 * L - 1), generator=gen, device=dev, dtype=torch.float64)
    flm = torch.complex(re, torch.where(mm == 0, torch.zeros_like(im), im))
    flm = torch.where((mm >= 0) & (mm <= ell), flm, torch.zeros_like(flm))
    sign = torch.where(mm % 2 == 0, 1.0, -1.0).to(torch.float64)
    flm = torch.where(mm < 0, torch.flip(flm, dims=[2]).conj() * sign, flm)
    f = sph.inverse_hip(flm, L, 0, nside, "healpix", True, plan=plan).real.contiguous()
    out = {}
    for name, pc in (("jacobi", None), ("learned", model)):
        hist, stats = [], {}
        refine.refine_mixed(f, L, 0, nside, "healpix", True, max_iter, plan=plan, precond=pc,
                            refresh_every=refresh_every, history=hist, flm_true=flm,
                            stats=stats, precision=precision)
        ok = [i for i, e in enumerate(hist) if e <= tol]
        out[name] = {"steps_to_tol": ok[0] if ok else None, "errs": hist, "stats": stats}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nside", type=int, default=256)
    ap.add_argument("--maps", type=int, default=4096)
    ap.add_argument("--keep", type=float, default=0.05)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--rows-per-step", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-nside", type=int, default=256)
    ap.add_argument("--out", default=str(lp.WEIGHTS))
    args = ap.parse_args()

    t0 = time.perf_counter()
    log(f"[train] data: {args.maps} maps at nside {args.nside}")
    X, Y = lp.make_training_rows(args.nside, args.maps, keep=args.keep, log=log)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    log(f"[train] {X.shape[0]} rows in {t_data:.1f} s")
    t0 = time.perf_counter()
    model, hist = lp.train(X, Y, steps=args.steps, rows_per_step=args.rows_per_step, lr=args.lr,
                           log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    model.save(args.out)
    log(f"[train] saved {args.out}")
    gain = eval_gain(model, args.eval_nside)
    print(json.dumps({"rows": int(X.shape[0]), "data_s": t_data, "train_s": t_train,
                      "loss": hist, "eval_nside": args.eval_nside, "gain": gain}), flush=True)


if __name__ == "__main__":
    main()
