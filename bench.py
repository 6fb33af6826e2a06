"""Headline benchmark: HEALPix forward transforms at L=2048, maps/s on one GPU.

Workload (one step): the forward spin-0 transform of a batch of 16 real
HEALPix maps at nside=1024 (L=2048), refined with Jacobi iterations
(flm += F(f - I(flm))) until the max-relative error against the seeded
coefficients is <= 1e-7.  Maps are band-limited: f = I(flm) for flm drawn
with seeds 300..315 (the reference's signal-generator distribution).  The
iteration count is fixed before timing by running the refinement once.

Prints one JSON line (rank 0) with the measured maps/s, the roofline of the
dominant kernel (timed live with HIP events on its launch stream) and a CPU
baseline from the NumPy oracle on a bounded sample of the same workload.
"""

import argparse
import json
import math
import os
import sys
import time

METRIC = "maps/s at L=2048 HEALPix: mixed+learned vs plain FP64 GPU at ≤1e-7 rel err"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nside", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--tol", type=float, default=1e-7)
    ap.add_argument("--max-iter", type=int, default=40)
    ap.add_argument("--seed0", type=int, default=300)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(nside, L, n_transforms, rings=256, msteps=256):
    """Oracle (NumPy, one core) on a bounded sample of the same workload.

    Sample: the oracle's Price-McEwen latitudinal recursion for `rings` of
    the 4*nside-1 rings and the first `msteps` of its L m-steps (every step
    costs the same: it updates full (rings x L) arrays), plus one full
    ring FFT of a map.  Scaled to one transform, times the transforms of a
    refined forward.
    """
    import numpy as np

    from oracle import ffts, geometry, pm

    ntheta = 4 * nside - 1
    beta = geometry.thetas(L, "healpix", nside)[:rings]
    rng = np.random.default_rng(0)
    ftm = rng.standard_normal((rings, 2 * L)) + 1j * rng.standard_normal((rings, 2 * L))
    t0 = time.perf_counter()
    gen = pm.dslices(beta, L, 0, True, 0, None, "healpix", nside, True)
    acc = 0.0
    for k, (m, D) in enumerate(gen):
        acc += float(np.nansum(D[:, -1:] * ftm[:, L + m : L + m + 1]).real)
        if k + 1 >= msteps:
            break
    t_lat = time.perf_counter() - t0
    f = rng.standard_normal(12 * nside * nside)
    t0 = time.perf_counter()
    ffts.healpix_fft(f, L, nside, True)
    t_fft = time.perf_counter() - t0
    per_transform = t_lat * (ntheta / rings) * (L / msteps) + t_fft
    per_map = per_transform * n_transforms
    return {
        "value": 1.0 / per_map,
        "unit": "maps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle PM recursion on {rings}/{ntheta} rings x {msteps}/{L} m-steps "
                   f"({t_lat:.1f} s) + one ring FFT ({t_fft:.1f} s), scaled to "
                   f"{n_transforms} transforms per map"),
    }


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from s2fft_amd import _lib
    from s2fft_amd.transforms import spherical as sph
    from s2fft_amd.utils.signal_generator import generate_flm

    nside, B = args.nside, args.batch
    L = 2 * nside
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    log(f"[bench] generating {B} maps nside={nside} L={L}")
    flm_np = np.stack([generate_flm(np.random.default_rng(args.seed0 + i), L, reality=True)
                       for i in range(B)])
    flm_true = torch.from_numpy(flm_np).to(dev)
    del flm_np
    plan = sph.get_plan("healpix", L, nside, 0, 0, True)
    f = sph.inverse_hip(flm_true, L, 0, nside, "healpix", True, plan=plan).real.contiguous()
    torch.cuda.synchronize()

    def max_rel(x):
        return float((x - flm_true).abs().max() / flm_true.abs().max())

    # fix the iteration count: refine until <= tol
    log("[bench] calibrating the refinement count")
    flm = sph.forward_hip(f, L, 0, nside, "healpix", True, plan=plan)
    errs = [max_rel(flm)]
    n_iter = 0
    while errs[-1] > args.tol and n_iter < args.max_iter:
        rec = sph.inverse_hip(flm, L, 0, nside, "healpix", True, plan=plan)
        flm = flm + sph.forward_hip(f - rec, L, 0, nside, "healpix", True, plan=plan)
        n_iter += 1
        errs.append(max_rel(flm))
        log(f"[bench]   iter {n_iter}: max-rel err {errs[-1]:.3e}")
    del flm, rec
    torch.cuda.empty_cache()

    def step():
        return sph.forward_iterative_hip(f, L, 0, nside, "healpix", True, 0, n_iter)

    for i in range(args.warmup):
        out = step()
        torch.cuda.synchronize()
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    final_err = max_rel(out) if args.warmup else float("nan")
    del out

    _lib.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"[bench] timed {args.steps} steps in {elapsed:.3f} s")
    lat_f = _lib.profile_query(_lib.KERNEL_LAT_FORWARD)
    lat_i = _lib.profile_query(_lib.KERNEL_LAT_INVERSE)
    fft_f = _lib.profile_query(_lib.KERNEL_RING_FFT_FORWARD)
    fft_i = _lib.profile_query(_lib.KERNEL_RING_FFT_INVERSE)
    _lib.profile_enable(False)
    if not args.warmup:
        final_err = max_rel(out)

    ntheta = 4 * nside - 1
    P = L * (L + 1) // 2
    flop_launch = math.ceil(ntheta / 2) * P * (5 + 4 * B)   # one launch = B maps
    kernels = {"lat_forward": lat_f, "lat_inverse": lat_i, "ring_fft_forward": fft_f,
               "ring_fft_inverse": fft_i}
    dom_name, (dom_ms, dom_n) = max(kernels.items(), key=lambda kv: kv[1][0])
    if dom_name.startswith("lat"):
        avg_ms = dom_ms / max(dom_n, 1)
        achieved = flop_launch / (avg_ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                "kernel": dom_name, "pipe": "fp64 VALU (peak equals the FP64 MFMA peak)",
                "flop_per_launch": flop_launch, "avg_launch_ms": avg_ms}
    else:
        avg_ms = dom_ms / max(dom_n, 1)
        nbytes = B * (8 * 12 * nside * nside + 16 * ntheta * L)
        achieved = nbytes / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                "frac": achieved / 8000.0, "traffic": None, "kernel": dom_name,
                "bytes_per_launch": nbytes, "avg_launch_ms": avg_ms}

    n_transforms = 1 + 2 * n_iter
    value = B * args.steps / elapsed
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"HEALPix nside={nside} L={L} real spin-0 forward of a {B}-map batch, "
                         f"FP64 Jacobi refinement to <= {args.tol:g} max-rel error"),
            "method": "plain FP64 HIP kernels",
            "nside": nside, "L": L, "batch": B, "refine_iters": n_iter,
            "transforms_per_step": n_transforms, "max_rel_err": final_err,
            "calibration_errs": errs,
            "kernel_ms": {k: {"total_ms": v[0], "launches": v[1]} for k, v in kernels.items()},
        },
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle sample)")
        result["cpu_baseline"] = cpu_baseline(nside, L, n_transforms)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
