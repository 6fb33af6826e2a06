"""Headline benchmark: HEALPix forward transforms at L=2048, maps/s on one GPU.

Workload (one step): the forward spin-0 transform of a batch of 16 real
HEALPix maps at nside=1024 (L=2048), refined until the max-relative error
against the seeded coefficients is <= 1e-7.  Maps are band-limited:
f = I(flm) for flm drawn with seeds 300..315 (the reference's
signal-generator distribution).  These pipelines are timed on the same maps:

* plain   -- the reference's Jacobi refinement flm += F(f - I(flm)) on the
             FP64 kernels, residual on the pixels (two ring FFTs per
             iteration): the "plain FP64 GPU" baseline of the metric;
* plain spectral -- the same FP64 iteration with the residual on the ring
             spectra (no ring FFT inside the loop);
* fp64 CG -- conjugate gradients on F I x = F f, every transform FP64: the
             algorithmic gain of CG alone;
* mixed CG + learned -- the headline (`value`): CG with the residual held in
             FP64 on the ring spectra, FP64 recursion, split-BF16 / BF16
             matrix-core projections and syntheses, one FP64 synthesis, and
             the learned MLP preconditioner on the first residual (its
             direction synthesised with the exact FP64 inverse and kept per
             map only where that FP64 ring residual drops);
* mixed CG -- the same without the network (the precision gain over fp64 CG);
* mixed Jacobi + learned -- Jacobi steps with the learned preconditioner on
             the first two, the FP64 residual refreshed every 2 steps;
* mixed Jacobi -- the same without the network, refreshed every 3 steps
             (each Jacobi leg at its fastest refresh period at nside 1024).

Step counts are fixed before timing by running each pipeline once against
the seeded coefficients, and confirmed on a second set of 16 maps (seeds
316..331): `config.legs[*].confirm_max_rel_err`.  Kernel times come from a
separate untimed pass (HIP events on the launch stream), never from the timed
steps.  A single-map leg runs the nside-512, L=1024 forward with iter=3 on the
FP64 kernels (`config.single_map`).  The CPU baseline is a torch-CPU twin of
the oracle's latitudinal recursion on all host threads, on a bounded sample.
"""

import argparse
import json
import math
import os
import sys
import time

METRIC = "maps/s at L=2048 HEALPix: mixed+learned vs plain FP64 GPU at ≤1e-7 rel err"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 matrix
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch at this workload (DRAM reads + writes) from rocprofv3 --pmc
# TCC_EA0_RDREQ_DRAM_32B_sum and TCC_EA0_WRREQ_WRITE_DRAM_32B_sum, 32 bytes each
# (request-size resolved: a 1 GiB device copy reads and writes 1.074e9 B by them):
# profiles/counters_healpix_nside1024_L2048_b16.md.  lat_forward: the FP64
# projection into the reference flm layout (the plain leg); lat_inverse: the FP64
# synthesis of the CG loop into tiled spectra; the BF16 kinds: the single-BF16
# projection and the split-BF16 synthesis of the CG loop (tiled spectra).
TRAFFIC_BYTES = {"lat_forward": 15.71e9, "lat_inverse": 3.16e9,
                 "lat_forward_bf16": 3.25e9, "lat_inverse_bf16": 4.88e9}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nside", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--tol", type=float, default=1e-7)
    ap.add_argument("--max-iter", type=int, default=40)
    ap.add_argument("--seed0", type=int, default=300)
    ap.add_argument("--confirm-seed0", type=int, default=316,
                    help="second seed set that confirms every leg's step count")
    ap.add_argument("--mixed-precision", default="bf16x1", choices=["bf16x1", "bf16", "mixed"])
    ap.add_argument("--mixed-inv-precision", default="bf16", choices=["bf16", "mixed"])
    ap.add_argument("--refresh-every", type=int, default=3,
                    help="Jacobi legs: FP64 residual refresh period without the network")
    ap.add_argument("--refresh-every-learned", type=int, default=2,
                    help="the same with the network (each leg at its fastest period)")
    ap.add_argument("--init-precision", default="bf16", choices=["fp64", "bf16", "mixed"])
    ap.add_argument("--learned-steps", type=int, default=2, help="Jacobi leg: learned steps")
    ap.add_argument("--cg-proj", default="bf16,bf16x1",
                    help="CG legs: projection arithmetic of the residuals r_0, r_1, ... "
                         "(the last repeats)")
    ap.add_argument("--cg-init", default="bf16", choices=["fp64", "bf16", "bf16x1", "mixed"])
    ap.add_argument("--cg-learned-steps", type=int, default=1)
    ap.add_argument("--no-jacobi-leg", action="store_true")
    ap.add_argument("--no-fp64-cg", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-confirm", action="store_true")
    ap.add_argument("--no-single-map", action="store_true")
    ap.add_argument("--single-map-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target duration of the CPU baseline sample")
    ap.add_argument("--kernels-once", action="store_true",
                    help="run each transform kind once at the bench workload, plus a 1 GiB "
                         "device copy, and exit (for rocprofv3 --pmc counter passes)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cpu():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nside, L, n_transforms, seconds):
    """Torch-CPU twin of the oracle's latitudinal m-recursion (the reference's
    JAX array program) on every host thread, on a bounded sample.

    Sample: `rings` equatorial-band rings of the 4 nside - 1 and the first
    `msteps` of the L m-steps (every step updates full (rings x L) arrays, so
    the time scales with both), timed as the minimum of 3 repeats, plus one
    full ring FFT of a map (torch.fft per ring); scaled to one transform, times
    the transforms of the plain refined forward.
    """
    import numpy as np
    import torch

    from oracle import geometry, torch_twin

    threads = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        threads = min(threads, int(env))
    torch.set_num_threads(threads)
    ntheta = 4 * nside - 1
    rings = 512
    beta = geometry.thetas(L, "healpix", nside)[nside:nside + rings]
    rng = np.random.default_rng(0)
    ftm = rng.standard_normal((rings, 2 * L)) + 1j * rng.standard_normal((rings, 2 * L))
    t0 = time.perf_counter()
    pre = torch_twin.precompute(beta, L)
    t_pre = time.perf_counter() - t0
    t0 = time.perf_counter()
    torch_twin.forward_latitudinal_sample(ftm, pre, L, 8)
    probe = (time.perf_counter() - t0) / 8
    msteps = int(max(8, min(L, seconds / 3 / max(probe, 1e-6))))
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        torch_twin.forward_latitudinal_sample(ftm, pre, L, msteps)
        best = min(best, time.perf_counter() - t0)
    # the reference recomputes its precomputes in every call (precomps=None)
    t_lat = (best * (L / msteps) + t_pre) * (ntheta / rings)
    # ring FFT of one map (real rfft per ring, torch)
    f = torch.from_numpy(rng.standard_normal(12 * nside * nside))
    lens = [4 * (t + 1) if t < nside - 1 else 4 * nside if t <= 3 * nside - 1
            else 4 * (4 * nside - t - 1) for t in range(ntheta)]
    t0 = time.perf_counter()
    off = 0
    for n in lens:
        torch.fft.rfft(f[off:off + n])
        off += n
    t_fft = time.perf_counter() - t0
    per_map = (t_lat + t_fft) * n_transforms
    return {
        "value": 1.0 / per_map,
        "unit": "maps/s",
        "cores": threads,
        "kind": "port",
        "host_cpu": host_cpu(),
        "sample": (f"torch-CPU twin of the latitudinal m-recursion on {rings}/{ntheta} rings x "
                   f"{msteps}/{L} m-steps, min of 3 repeats ({best:.2f} s), its precomputes "
                   f"for those rings ({t_pre:.2f} s), plus one ring FFT "
                   f"({t_fft:.2f} s), scaled to {n_transforms} transforms per map "
                   f"on {threads} threads"),
    }


KERNEL_NAMES = ("lat_forward", "lat_inverse", "lat_forward_fp32", "lat_inverse_fp32",
                "lat_forward_bf16", "lat_inverse_bf16", "ring_fft_forward", "ring_fft_inverse",
                "ring_residual")


def kernel_times(_lib):
    """(total ms, launches) per kernel kind; the latitudinal kinds by arithmetic."""
    codes = {"lat_forward": _lib.KERNEL_LAT_FORWARD, "lat_inverse": _lib.KERNEL_LAT_INVERSE,
             "lat_forward_fp32": _lib.KERNEL_LAT_FORWARD_FP32,
             "lat_inverse_fp32": _lib.KERNEL_LAT_INVERSE_FP32,
             "lat_forward_bf16": _lib.KERNEL_LAT_FORWARD_BF16,
             "lat_inverse_bf16": _lib.KERNEL_LAT_INVERSE_BF16,
             "ring_fft_forward": _lib.KERNEL_RING_FFT_FORWARD,
             "ring_fft_inverse": _lib.KERNEL_RING_FFT_INVERSE,
             "ring_residual": _lib.KERNEL_RING_RESIDUAL}
    out = {k: _lib.profile_query(v) for k, v in codes.items()}
    return {k: v for k, v in out.items() if v[1] > 0}


def mfma_f64_ceiling(_lib, torch, blocks=4096, iters=250):
    """Sustained V_MFMA_F64_16X16X4F64 TFLOP/s of this device (s2h_probe_mfma_f64_rate)."""
    out = torch.empty(blocks * 256, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream()
    _lib.check(_lib.lib().s2h_probe_mfma_f64_rate(blocks, 10, out.data_ptr(), s.cuda_stream))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    _lib.check(_lib.lib().s2h_probe_mfma_f64_rate(blocks, iters, out.data_ptr(), s.cuda_stream))
    e1.record(s)
    e1.synchronize()
    ms = e0.elapsed_time(e1)
    return blocks * 4 * iters * 8 * 2048.0 / (ms * 1e-3) / 1e12


def lat_flop(nside, L, B):
    """Algorithmic flops of one latitudinal launch: ceil(ntheta/2) ring pairs x
    P = L(L+1)/2 coefficients x (5 + 4 B)."""
    return math.ceil((4 * nside - 1) / 2) * (L * (L + 1) // 2) * (5 + 4 * B)


def roofline(kernels, nside, L, B):
    """Roofline entry of the dominant kernel kind (by total time).

    Latitudinal kernels: algorithmic flops per launch (lat_flop), priced
    against the peak of the launch's arithmetic (FP64, FP32 or dense BF16).
    Ring FFTs: algorithmic bytes per launch = B maps x (8 B per pixel + 16 B
    per ftm element, ntheta x L).
    """
    dom, (ms, n) = max(kernels.items(), key=lambda kv: kv[1][0])
    avg_ms = ms / max(n, 1)
    ntheta = 4 * nside - 1
    if dom.startswith("lat"):
        flop = lat_flop(nside, L, B)
        achieved = flop / (avg_ms * 1e-3) / 1e12
        arith = "bf16" if dom.endswith("bf16") else "fp32" if dom.endswith("fp32") else "fp64"
        peak = {"bf16": BF16_PEAK_TFLOPS, "fp32": FP32_PEAK_TFLOPS}.get(arith, FP64_PEAK_TFLOPS)
        return {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": TRAFFIC_BYTES.get(dom), "kernel": dom,
                "arith": arith, "flop_per_launch": flop, "avg_launch_ms": avg_ms}
    nbytes = B * (8 * 12 * nside * nside + 16 * ntheta * L)
    achieved = nbytes / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_BYTES.get(dom), "kernel": dom,
            "bytes_per_launch": nbytes, "avg_launch_ms": avg_ms}


def kernels_once(nside, B):
    """One launch of each transform kind at the bench workload, for counters."""
    import torch

    from s2fft_amd.learned.precond import MLPPreconditioner
    from s2fft_amd.transforms import refine
    from s2fft_amd.transforms import spherical as sph

    L = 2 * nside
    dev = torch.device("cuda:0")
    plan = sph.get_plan("healpix", L, nside, 0, 0, True, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    f = torch.randn((B, 12 * nside * nside), generator=g, dtype=torch.float64, device=dev)
    flm = sph.forward_hip(f, L, 0, nside, "healpix", True, plan=plan)
    sph.inverse_hip(flm, L, 0, nside, "healpix", True, plan=plan)
    for prec in ("mixed", "bf16", "bf16x1"):
        sph.forward_hip(f, L, 0, nside, "healpix", True, plan=plan, precision=prec)
    for prec in ("mixed", "bf16", "bf16x1"):
        sph.inverse_hip(flm, L, 0, nside, "healpix", True, plan=plan, precision=prec)
    pc = MLPPreconditioner.load(device=dev)
    pc.apply(flm, L, True, 0)
    G = sph.ring_analysis_hip(f, plan)
    R = sph.ring_residual_hip(G, sph.inverse_rings_hip(flm, plan), plan)
    sph.forward_rings_hip(R, L, plan, "bf16x1")
    sph.inverse_hip(flm, L, 0, nside, "healpix", True, plan=plan, real_output=True)
    refine.refine_cg(f, L, 0, nside, "healpix", True, 3, plan=plan, precond=pc, learned_steps=1)
    # the single-map workload: nside 512, one real map, forward with iter=3
    # (the VALU latitudinal kernels)
    n1 = 512
    f1 = torch.randn((1, 12 * n1 * n1), generator=g, dtype=torch.float64, device=dev)
    sph.forward_iterative_hip(f1, 2 * n1, 0, n1, "healpix", True, 0, 3)
    src = torch.empty(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
    src.fill_(1.0)
    torch.empty_like(src).copy_(src)
    torch.cuda.synchronize()
    print("bench.py --kernels-once: done", flush=True)


def single_map_leg(args, torch, _lib, sph, generate_flm):
    """nside 512, L 1024, one real map (seed 1): forward with iter=3 on the
    FP64 kernels (one map: the VALU latitudinal kernels), maps/s and the
    kernels against their rooflines."""
    import numpy as np

    nside, L = 512, 1024
    flm = torch.from_numpy(generate_flm(np.random.default_rng(1), L, reality=True)).cuda()
    f = sph.inverse_hip(flm[None], L, 0, nside, "healpix", True).real.contiguous()
    run = lambda: sph.forward_iterative_hip(f, L, 0, nside, "healpix", True, 0, 3)
    out = None
    for _ in range(max(args.warmup, 1)):
        out = run()
    torch.cuda.synchronize()
    err = float((out[0] - flm).abs().max() / flm.abs().max())
    t0 = time.perf_counter()
    for _ in range(args.single_map_steps):
        run()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    _lib.profile_enable(True)
    run()
    torch.cuda.synchronize()
    kt = kernel_times(_lib)
    _lib.profile_enable(False)
    res = {"workload": "HEALPix nside=512 L=1024 real spin-0 forward, iter=3, one map, FP64",
           "maps_per_s": args.single_map_steps / el, "ms_per_map": 1e3 * el / args.single_map_steps,
           "max_rel_err_vs_flm": err,
           "kernel_ms": {k: {"total_ms": v[0], "launches": v[1]} for k, v in kt.items()}}
    flop = lat_flop(nside, L, 1)
    for k in ("lat_forward", "lat_inverse"):
        if k in kt:
            ms = kt[k][0] / kt[k][1]
            res[k + "_fp64_tflops"] = flop / (ms * 1e-3) / 1e12
            res[k + "_frac_of_fp64_peak"] = res[k + "_fp64_tflops"] / FP64_PEAK_TFLOPS
    if "ring_fft_forward" in kt:
        ms = kt["ring_fft_forward"][0] / kt["ring_fft_forward"][1]
        nbytes = 8 * 12 * nside * nside + 16 * (4 * nside - 1) * L
        res["ring_fft_forward_gbs"] = nbytes / (ms * 1e-3) / 1e9
    return res


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from s2fft_amd import _lib
    from s2fft_amd.transforms import refine
    from s2fft_amd.transforms import spherical as sph
    from s2fft_amd.utils.signal_generator import generate_flm

    nside, B = args.nside, args.batch
    L = 2 * nside
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.kernels_once:
        kernels_once(nside, B)
        return

    plan = sph.get_plan("healpix", L, nside, 0, 0, True)

    def make_maps(seed0):
        flm_np = np.stack([generate_flm(np.random.default_rng(seed0 + i), L, reality=True)
                           for i in range(B)])
        flm_t = torch.from_numpy(flm_np).to(dev)
        del flm_np
        fm = sph.inverse_hip(flm_t, L, 0, nside, "healpix", True, plan=plan).real.contiguous()
        torch.cuda.synchronize()
        return flm_t, fm

    log(f"[bench] generating {B} maps nside={nside} L={L}")
    flm_true, f = make_maps(args.seed0)
    confirm = None
    if not args.no_confirm:
        log(f"[bench] generating the confirmation set (seeds {args.confirm_seed0}..)")
        confirm = make_maps(args.confirm_seed0)

    def max_rel(x, ref):
        return float((x - ref).abs().max() / ref.abs().max())

    def timed(step, label):
        """W warmups, K timed steps (no profiling), then one untimed profiled
        step for the kernel times (per step)."""
        out = None
        for i in range(args.warmup):
            out = step()
            torch.cuda.synchronize()
            log(f"[bench] {label} warmup {i + 1}/{args.warmup} done")
        err = max_rel(out, flm_true) if out is not None else float("nan")
        del out
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = step()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        if not args.warmup:
            err = max_rel(out, flm_true)
        del out
        _lib.profile_enable(True)
        step()
        torch.cuda.synchronize()
        kt = kernel_times(_lib)
        _lib.profile_enable(False)
        log(f"[bench] {label}: {args.steps} steps in {elapsed:.3f} s, err {err:.2e}")
        return elapsed, err, kt

    def confirm_err(run):
        if confirm is None:
            return None
        flm2, f2 = confirm
        return max_rel(run(f2), flm2)

    legs = {}
    # ---------------------------------------------------------------- plain
    plain = None
    if not args.no_plain:
        log("[bench] plain FP64: calibrating the Jacobi count")
        flm = sph.forward_hip(f, L, 0, nside, "healpix", True, plan=plan)
        errs = [max_rel(flm, flm_true)]
        n_plain = 0
        while errs[-1] > args.tol and n_plain < args.max_iter:
            rec = sph.inverse_hip(flm, L, 0, nside, "healpix", True, plan=plan)
            flm = flm + sph.forward_hip(f - rec, L, 0, nside, "healpix", True, plan=plan)
            n_plain += 1
            errs.append(max_rel(flm, flm_true))
            log(f"[bench]   iter {n_plain}: max-rel err {errs[-1]:.3e}")
        del flm, rec
        torch.cuda.empty_cache()
        pix = lambda fm: refine.refine_fp64(fm, L, 0, nside, "healpix", True, n_plain)
        el, err, kt = timed(lambda: pix(f), "plain")
        plain = {"elapsed": el, "err": err, "kernels": kt, "iters": n_plain, "calib": errs,
                 "roofline": roofline(kt, nside, L, B), "confirm": confirm_err(pix)}
        spec = lambda fm: refine.refine_fp64(fm, L, 0, nside, "healpix", True, n_plain,
                                             spectral=True)
        el, err, kt = timed(lambda: spec(f), "plain spectral")
        legs["plain_fp64_spectral"] = {"elapsed": el, "err": err, "kernels": kt,
                                       "iters": n_plain, "calib": errs, "stats": {},
                                       "roofline": roofline(kt, nside, L, B),
                                       "confirm": confirm_err(spec)}

    # ---------------------------------------------------------------- CG legs
    precond, learned = None, False
    if not args.no_learned:
        from s2fft_amd.learned import precond as lp
        if lp.WEIGHTS.exists():
            precond = lp.MLPPreconditioner.load(device=dev)
            learned = True
        else:
            log("[bench] no trained preconditioner weights: the mixed legs run without it")
    cg_proj = tuple(args.cg_proj.split(","))

    def calibrate(label, run):
        hist, stats = [], {}
        run(f, args.max_iter, hist, stats)
        ok = [i for i, e in enumerate(hist) if e <= args.tol]
        n = ok[0] if ok else args.max_iter
        log(f"[bench]   {label} errors {['%.1e' % e for e in hist[:n + 2]]} -> {n} steps")
        torch.cuda.empty_cache()
        return n, hist, stats

    def cg_run(pc, lsteps, fp64=False):
        def run(fm, n, hist=None, stats=None):
            kw = dict(init_precision="fp64", proj_precision=("fp64",), synth_precision="fp64",
                      last_synth_precision="fp64") if fp64 else \
                dict(init_precision=args.cg_init, proj_precision=cg_proj, synth_precision="bf16")
            return refine.refine_cg(fm, L, 0, nside, "healpix", True, n, plan=plan, precond=pc,
                                    learned_steps=lsteps, history=hist,
                                    flm_true=flm_true if hist is not None else None,
                                    stats=stats, **kw)
        return run

    def jacobi_run(pc):
        rf = args.refresh_every_learned if pc is not None else args.refresh_every

        def run(fm, n, hist=None, stats=None):
            return refine.refine_mixed(
                fm, L, 0, nside, "healpix", True, n, plan=plan, precond=pc,
                refresh_every=rf, history=hist,
                flm_true=flm_true if hist is not None else None, stats=stats,
                precision=args.mixed_precision, inv_precision=args.mixed_inv_precision,
                learned_steps=args.learned_steps, spectral=True,
                init_precision=args.init_precision)
        return run

    specs = [("mixed_cg_learned", cg_run(precond, args.cg_learned_steps))] if learned else []
    specs.append(("mixed_cg", cg_run(None, 0)))
    if not args.no_fp64_cg:
        specs.append(("fp64_cg", cg_run(None, 0, fp64=True)))
    if not args.no_jacobi_leg:
        # the same Jacobi refinement with and without the network: where the
        # network saves steps
        if learned:
            specs.append(("mixed_jacobi_learned", jacobi_run(precond)))
        specs.append(("mixed_jacobi", jacobi_run(None)))
    for label, run in specs:
        log(f"[bench] {label}: calibrating the step count")
        n, hist, stats = calibrate(label, run)
        el, err, kt = timed(lambda: run(f, n), label)
        legs[label] = {"elapsed": el, "err": err, "kernels": kt, "iters": n, "calib": hist[:n + 1],
                       "stats": stats, "roofline": roofline(kt, nside, L, B),
                       "confirm": confirm_err(lambda fm: run(fm, n))}
    head_label = "mixed_cg_learned" if learned else "mixed_cg"
    mixed = legs[head_label]

    ceiling = mfma_f64_ceiling(_lib, torch)
    log(f"[bench] FP64 matrix-core rate probe: {ceiling:.1f} TFLOP/s")
    for leg in ([plain] if plain else []) + list(legs.values()):
        rf = leg["roofline"]
        if rf.get("arith") == "fp64":
            rf["measured_fp64_mfma_tflops"] = ceiling
            rf["frac_of_measured_fp64_mfma"] = rf["achieved"] / ceiling
    value = B * args.steps / mixed["elapsed"]
    plain_value = B * args.steps / plain["elapsed"] if plain else None
    # the dominant kernel of the measured (headline) leg
    head = mixed["roofline"]

    def leg_json(v):
        d = {"maps_per_s": B * args.steps / v["elapsed"],
             "ms_per_step": 1e3 * v["elapsed"] / args.steps, "steps": v["iters"],
             "max_rel_err": v["err"], "confirm_max_rel_err": v.get("confirm"),
             "calibration_errs": v["calib"],
             "learned_accepted": v["stats"].get("accepted"),
             "learned_judged": v["stats"].get("steps"),
             "kernel_ms_per_step": {kk: {"total_ms": vv[0], "launches": vv[1]}
                                    for kk, vv in v["kernels"].items()},
             "roofline": v["roofline"]}
        if plain_value:
            d["speedup_vs_plain_fp64"] = d["maps_per_s"] / plain_value
        return d

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "maps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mixed["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"HEALPix nside={nside} L={L} real spin-0 forward of a {B}-map batch "
                         f"refined to <= {args.tol:g} max-rel error vs the seeded flm"),
            "method": (f"mixed: conjugate gradients on F I x = F f with the residual held in "
                       f"FP64 on the ring spectra (no ring FFT in the loop); FP64 recursion, "
                       f"{args.cg_init} first projection, {'/'.join(cg_proj)} matrix-core "
                       f"projections of the residuals, split-BF16 matrix-core syntheses of the "
                       f"directions (the last single BF16), one FP64 synthesis; spectra in "
                       f"blocks of 8 rings, the CG dot fused into the syntheses"
                       + (f"; learned MLP preconditioner on the first {args.cg_learned_steps} "
                          f"residual(s), its direction synthesised with the exact FP64 inverse "
                          f"and kept per map only where that FP64 ring residual drops"
                          if learned else "")),
            "nside": nside, "L": L, "batch": B, "learned": learned,
            "mixed_steps": mixed["iters"], "mixed_max_rel_err": mixed["err"],
            "mixed_confirm_max_rel_err": mixed.get("confirm"),
            "confirm_seeds": (f"{args.confirm_seed0}..{args.confirm_seed0 + B - 1}"
                              if confirm is not None else None),
            "mixed_roofline": mixed["roofline"],
            "legs": {k: leg_json(v) for k, v in legs.items()},
        },
        "roofline": head,
    }
    if plain:
        result["config"].update({
            "plain_roofline": plain["roofline"],
            "plain_fp64_maps_per_s": plain_value,
            "speedup_vs_plain_fp64": value / plain_value,
            "plain_ms_per_step": 1e3 * plain["elapsed"] / args.steps,
            "plain_refine_iters": plain["iters"], "plain_max_rel_err": plain["err"],
            "plain_confirm_max_rel_err": plain["confirm"],
            "plain_calibration_errs": plain["calib"],
            "plain_kernel_ms_per_step": {k: {"total_ms": v[0], "launches": v[1]}
                                         for k, v in plain["kernels"].items()},
        })
    del f, flm_true, confirm
    torch.cuda.empty_cache()
    if not args.no_single_map:
        log("[bench] single-map leg: nside 512, iter=3")
        result["config"]["single_map"] = single_map_leg(args, torch, _lib, sph, generate_flm)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (torch-CPU twin of the oracle recursion)")
        n_tr = 1 + 2 * (plain["iters"] if plain else mixed["iters"])
        result["cpu_baseline"] = cpu_baseline(nside, L, n_tr, args.cpu_seconds)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
