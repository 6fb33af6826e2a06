"""Iterative refinement of the forward transform on the GPU.

The reference refines a HEALPix forward transform with Jacobi iterations
flm <- flm + F(f - I(flm)) (s2fft/utils/iterative_refinement.py:9-46, called
from s2fft/transforms/spherical.py:453-462), all in FP64.  ``refine_fp64``
is that iteration on the FP64 kernels.  ``refine_mixed`` keeps the residual
in FP64 but runs every projection and synthesis on the mixed-precision
kernels:

    flm = F~(f);  r = f - I(flm)              (FP64 synthesis: exact residual)
    repeat:  h = F~(r);  delta = P(h);  flm += delta;  r -= I~(delta)

F~ / I~ are the FP32 matrix-core transforms (relative error eps ~ 1e-7).
The update delta is a correction whose error is eps |delta|, so the error
contracts by rho + eps per step (rho ~ 0.13 for HEALPix at L = 2 nside);
the incremental residual carries an error below eps sum_k |I delta_k|, which
is about eps rho / (1 - rho) of |f| because the first residual is exact.
P is the identity (Jacobi) or a learned per-(l, m) preconditioner whose step
is kept only where the residual, measured with the exact FP64 synthesis of
the step, decreases.

With ``spectral`` (plans with one ring grid: HEALPix) the residual lives on
the ring spectra instead of the pixels: r = G - A(Lat_inv(flm)) with
G the ring FFT of f (computed once) and A = (ring FFT) . (ring iFFT) applied
without either FFT (s2h_ring_residual), so F~(r) = Lat_fwd~(r) and the loop
runs no ring FFT at all.  The iterates are the same up to rounding.  With
real maps the coefficients stay in the packed m-major layout inside the loop
(the l of one m contiguous: whole cache lines for the latitudinal kernels,
half the memory) and are unpacked once at the end.
"""

import torch

from .. import _lib
from . import spherical as sph


def refine_fp64(f, L, spin, nside, sampling, reality, n_iter, L_lower=0, plan=None,
                spectral=False):
    """The reference's Jacobi refinement on the FP64 kernels (n_iter steps);
    ``spectral``: residuals on the ring spectra (no ring FFT per step)."""
    return sph.forward_iterative_hip(f, L, spin, nside, sampling, reality, L_lower, n_iter,
                                     spectral=spectral)


def refine_mixed(f, L, spin, nside, sampling, reality, n_iter, L_lower=0, plan=None,
                 precond=None, refresh_every=0, history=None, flm_true=None, stats=None,
                 precision="mixed", inv_precision=None, learned_steps=None, spectral=False,
                 init_precision="fp64", judge_precision="fp64"):
    """Mixed-precision refinement with an FP64 residual (see module docstring).

    precond: None (Jacobi) or an object with ``apply(h, L, reality, lmin)``
    returning a correction (the learned preconditioner); its steps are judged
    per map against the residual (``_judged_step``).  learned_steps: how many
    of the first steps use it (default all but the last, whose residual is
    never formed); later steps are Jacobi.
    refresh_every: recompute the residual with the FP64 synthesis every this
    many steps (0 = only at the start).
    history: optional list that receives the max-relative error per step when
    ``flm_true`` is given.  stats: optional dict counting accepted learned steps.
    precision / inv_precision: arithmetic of F~ / I~ ("mixed" FP32 or "bf16"
    split-BF16 matrix cores; inv_precision defaults to precision).
    spectral: residuals on the ring spectra (see the module docstring); needs
    a plan with one ring grid.
    init_precision: arithmetic of the synthesis that forms the first residual
    (FP64 by default; a reduced one leaves its error in r until a refresh).
    judge_precision: arithmetic of the synthesis that measures a learned
    step's residual (the exact FP64 inverse by default).
    """
    s = sampling.lower()
    plan = plan or sph.get_plan(s, L, nside, spin, L_lower, reality, device=f.device)
    kw = dict(L=L, spin=spin, nside=nside, sampling=s, reality=reality, L_lower=L_lower,
              plan=plan)
    fp, ip = precision, inv_precision or precision
    if learned_steps is None:
        learned_steps = max(n_iter - 1, 0)
    real_out = bool(reality) and not f.is_complex()
    packed = False
    if spectral:
        # f is replaced by its ring spectra G: every residual is G - A(I~(x))
        f = sph.ring_analysis_hip(f, plan)
        packed = bool(reality)

        def synth(x, prec="fp64", sub=None):
            return sph.ring_residual_hip(sub, sph.inverse_rings_hip(x, plan, prec, packed), plan)

        def proj(x, prec, acc=None):
            return sph.forward_rings_hip(x, L, plan, prec, accumulate_into=acc, packed=packed)
    else:
        def synth(x, prec="fp64", sub=None):
            return sph.inverse_hip(x, precision=prec, real_output=real_out, sub_from=sub, **kw)

        def proj(x, prec, acc=None):
            return sph.forward_hip(x, precision=prec, accumulate_into=acc, **kw)

    unpack = (lambda x: sph.flm_unpack_hip(x, L)) if packed else (lambda x: x)
    flm = proj(f, fp)
    r = synth(flm, init_precision, sub=f)
    rn = None  # per-map squared norm of r, when known
    _track(history, flm, flm_true, unpack)
    for k in range(n_iter):
        last = k + 1 == n_iter
        if last:
            proj(r, fp, acc=flm)
            _track(history, flm, flm_true, unpack)
            break
        h = proj(r, fp)
        refresh = bool(refresh_every and (k + 1) % refresh_every == 0)
        if precond is not None and k < learned_steps:
            delta, r, rn = _judged_step(precond, h, r, rn, f, flm, kw, stats, synth, ip, refresh,
                                        packed, judge_precision)
            flm += delta
        else:
            flm += h
            r = synth(flm, sub=f) if refresh else synth(h, ip, sub=r)
            rn = None
        _track(history, flm, flm_true, unpack)
    return unpack(flm)


def refine_cg(f, L, spin, nside, sampling, reality, n_iter, L_lower=0, plan=None, precond=None,
              learned_steps=0, init_precision="bf16", proj_precision=("bf16", "bf16x1"),
              synth_precision="bf16", last_synth_precision="bf16x1", history=None,
              flm_true=None, stats=None, judge_precision="fp64"):
    """Conjugate-gradient refinement of the forward transform on the ring spectra.

    F I is Hermitian positive definite on the coefficients (the HEALPix and
    GL forwards are the quadrature-weighted adjoints of the inverses), so
    flm solves F I flm = F f, and CG reaches a given error in fewer
    transform pairs than the reference's Jacobi iteration
    (s2fft/utils/iterative_refinement.py:9-46).  With G the ring spectra of f:

        x_0 = F~(G);  R = G - A(I(x_0))            (FP64 synthesis)
        r_0 = F~(R);  z_0 = P(r_0) or r_0;  p_0 = z_0
        step k:  Y = I~(p_k);  pq_k = <p_k, F A(Y)>  (a ring-space sum)
                 alpha_k = <r_k, z_k> / pq_k;  x += alpha_k p_k
                 R -= alpha_k A(Y);  r_{k+1} = F~(R);  z = P(r_{k+1}) or r_{k+1}
                 p_{k+1} = z - sum_j <z, q_j> / pq_j p_j   (j = k, and k - 1
                 when p_{k-1} was a learned direction),
                 q_j = F I p_j = (r_j - r_{j+1}) / alpha_j

    R stays FP64 and takes each direction's synthesis, so every r_k is the
    projection of a residual that is accurate to the reduced synthesis' error
    times |alpha p| (not a running sum of reduced-precision products).  A
    direction that follows a learned one is orthogonalised against both
    (flexible CG, for a nonlinear P).  The last step forms no residual.

    F~ / I~ run at proj_precision (one entry per residual r_0, r_1, ...; the
    last repeats) / synth_precision; x_0 at init_precision.  The last step's
    synthesis only sets its step length (no residual is formed from it), so
    it runs at last_synth_precision.  precond: the
    learned preconditioner for the first ``learned_steps`` residuals; a
    learned direction is synthesised at judge_precision (the exact FP64
    inverse by default), so the residual R that takes its step is the FP64
    residual of the new iterate, and it is kept for a map only where that
    step lowers the norm of R, else the map redoes the step along its CG
    direction (stats counts accepted / judged).  Needs a plan with one ring
    grid (HEALPix, GL).  Returns flm (B, L, 2L-1).
    """
    s = sampling.lower()
    plan = plan or sph.get_plan(s, L, nside, spin, L_lower, reality, device=f.device)
    if not sph.spectra_supported(plan):
        raise ValueError("refine_cg needs a sampling whose forward and inverse share one ring grid")
    packed = bool(reality)
    # real maps: R and every Y in the tiled spectra layout (the latitudinal
    # kernels' column accesses then touch whole cache lines); G stays as the
    # ring analysis writes it
    tl = packed
    pL = L if packed else 0
    unpack = (lambda x: sph.flm_unpack_hip(x, L)) if packed else (lambda x: x)
    lmin = max(abs(spin), L_lower)

    def proj(R, k):
        prec = proj_precision[min(k, len(proj_precision) - 1)]
        return sph.forward_rings_hip(R, L, plan, prec, packed=packed, tiled=tl)

    def precondition(r, k):
        if precond is not None and k < learned_steps:
            return precond.apply(r, L, reality, lmin, packed=packed)
        return r

    G = sph.ring_analysis_hip(f, plan)
    x = sph.forward_rings_hip(G, L, plan, init_precision, packed=packed)
    _track(history, x, flm_true, unpack)
    if n_iter <= 0:
        return unpack(x)
    def judged(k):
        return precond is not None and k < learned_steps

    # R = G - A(I(x_0)) (tiled: into a new array, else in place of G); |R|^2
    # per map where a judged step follows
    Y = sph.inverse_rings_hip(x, plan, "fp64", packed, tiled=tl)
    n_R = None
    kw_r = dict(out=None, tiled=True, g_tiled=False) if tl else dict(out=G)
    if judged(0):
        R, n_R = sph.ring_residual_hip(G, Y, plan, sumsq=True, **kw_r)
    else:
        R = sph.ring_residual_hip(G, Y, plan, **kw_r)
    del G
    r = proj(R, 0)
    if judged(0) and hasattr(precond, "apply_dot"):
        # the network's pass also forms <z_0, r_0>
        z, rz = precond.apply_dot(r, L, reality, lmin, packed=packed)
    else:
        z = precondition(r, 0)
        rz = sph.flm_dots_hip(z, [r], pL)[0]
    # p_0 may be r_0 itself: no direction buffer is written before r_0 is dropped
    rs, ps, rz_prev = [r], [z], None
    # per map: whether direction p_{k-1} / p_k is a kept learned direction
    kept_prev = kept = None
    # x += alpha_k p_k is deferred (one pass over x for several directions,
    # fused with the unpack at the end) unless every iterate is tracked
    eager = history is not None
    pending = []
    for k in range(n_iter):
        p = ps[-1]
        last = k + 1 == n_iter
        prec = judge_precision if judged(k) else \
            last_synth_precision if last else synth_precision
        if tl:
            Y, pq, alpha = sph.inverse_rings_dot_hip(p, plan, prec, packed, num=rz)
        else:
            Y = sph.inverse_rings_hip(p, plan, prec, packed)
            pq, alpha = sph.ring_dot_hip(Y, plan, num=rz)
        if last:
            if eager:
                sph.cg_update_hip(x, p, alpha)
                _track(history, x, flm_true, unpack)
                break
            pending.append((p, alpha))
            return sph.cg_finish_hip(x, [q for q, _ in pending], [a for _, a in pending],
                                     unpack_L=L if packed else None)
        n_new = None
        if judged(k) or judged(k + 1):
            _, n_new = sph.ring_residual_hip(R, Y, plan, out=R, alpha=alpha, sumsq=True,
                                             tiled=tl)
        else:
            sph.ring_residual_hip(R, Y, plan, out=R, alpha=alpha, tiled=tl)
        if judged(k):
            ok = n_new < n_R
            if stats is not None:
                stats["accepted"] = stats.get("accepted", 0) + int(ok.sum())
                stats["steps"] = stats.get("steps", 0) + int(ok.numel())
            if not bool(ok.all()):
                _redo_plain((~ok).nonzero().flatten(), R, Y, p, alpha, rz, rs[-1], plan,
                            synth_precision, packed, pL, n_new, tl)
            kept = ok
        n_R = n_new
        r_new = proj(R, k + 1)
        z = precondition(r_new, k + 1)
        # d = <z, r_k>, <z, r_{k+1}> [, <z, r_{k-1}>]; alpha_j pq_j = <r_j, z_j>.
        # p_{k-1} enters only when it was a learned direction: with z = r the
        # two-term recurrence already keeps it conjugate, and the extra term
        # would add the reduced projections' error of r_{k-1} (at single BF16
        # it costs an order of magnitude by the third step).  Maps whose learned
        # direction was rejected get g2 = 0 (no host synchronisation to test
        # whether any was kept).
        back2 = len(rs) > 1 and kept_prev is not None
        d = sph.flm_dots_hip(z, [rs[-1], r_new] + (rs[-2:-1] if back2 else []), pL)
        g1 = (d[0] - d[1]) / rz
        g2 = (d[2] - d[0]) / rz_prev * kept_prev if back2 else None
        p_new = ps[0] if len(ps) == 3 else torch.empty_like(p)
        if not eager and len(ps) == 3:
            # ps[0] is reused below: its pending update goes into x first
            sph.cg_finish_hip(x, [q for q, _ in pending], [a for _, a in pending])
            pending = []
        sph.cg_update_hip(x if eager else None, p, alpha, p_new, z, g1,
                          ps[-2] if back2 else None, g2)
        if not eager:
            pending.append((p, alpha))
        _track(history, x, flm_true, unpack)
        kept_prev, kept = kept, None
        rz_prev, rz = rz, d[1]
        rs = rs[-1:] + [r_new]
        ps = ps[-2:] + [p_new]
    return unpack(x)


def _redo_plain(bad, R, Y, p, alpha, rz, r, plan, synth_precision, packed, pL, n_R, tl=False):
    """Judge's fallback of refine_cg: maps ``bad`` take back their learned
    step's residual update (R += alpha A(Y)) and step along r instead; R and
    the maps' rows of p, alpha, rz and the residual norms n_R are replaced in
    place (x is updated afterwards with the replaced alpha and p).  ``tl``:
    R and Y in the tiled spectra layout."""
    Rb = R[bad].contiguous()
    sph.ring_residual_hip(Rb, Y[bad].contiguous(), plan, out=Rb, alpha=-alpha[bad], tiled=tl)
    rb = r[bad].contiguous()
    rzb = sph.flm_dots_hip(rb, [rb], pL)[0]
    if tl:
        Yb, _, ab = sph.inverse_rings_dot_hip(rb, plan, synth_precision, packed, num=rzb)
    else:
        Yb = sph.inverse_rings_hip(rb, plan, synth_precision, packed)
        _, ab = sph.ring_dot_hip(Yb, plan, num=rzb)
    _, nb = sph.ring_residual_hip(Rb, Yb, plan, out=Rb, alpha=ab, sumsq=True, tiled=tl)
    R[bad], p[bad] = Rb, rb
    alpha[bad], rz[bad], n_R[bad] = ab, rzb, nb


def _judged_step(precond, h, r, rn, f, flm, kw, stats, synth, ip, exact, packed=False,
                 judge_precision="fp64"):
    """Learned step per map where it lowers the FP64-held residual, else Jacobi.

    The candidate delta = P(h) moves the residual to r - I(delta), with I the
    synthesis at judge_precision (the exact FP64 inverse by default; on a
    refresh step the whole residual f - I(flm + delta) is re-formed in FP64);
    a map keeps it only if that residual's 2-norm is below the current one
    (rn: the current squared norms, or None).  Rejected maps take the Jacobi
    step delta = h instead (one more synthesis at the reduced precision ip,
    for those maps only).
    Returns (delta, new residual, its squared norms where known, else None).
    """
    L = kw["L"]
    lmin = max(abs(kw["spin"]), kw["L_lower"])
    delta = precond.apply(h, L, kw["reality"], lmin, packed=packed)

    def residual(d, rows, prec):
        if exact:
            return synth((flm[rows] + d).contiguous(), sub=f[rows])
        return synth(d.contiguous(), prec, sub=r[rows])

    every = slice(None)
    r_new = residual(delta, every, judge_precision)
    new_n = _sumsq(r_new)
    ok = new_n < (_sumsq(r) if rn is None else rn)
    if stats is not None:
        stats["accepted"] = stats.get("accepted", 0) + int(ok.sum())
        stats["steps"] = stats.get("steps", 0) + int(ok.numel())
    if bool(ok.all()):
        return delta, r_new, new_n
    bad = (~ok).nonzero().flatten()
    hj = h[bad].contiguous()
    rj = residual(hj, bad, ip)
    delta = delta.clone()
    r_new = r_new.clone()
    delta[bad] = hj
    r_new[bad] = rj
    return delta, r_new, None


def _sumsq(x):
    """Per-map sum of squares on the device (s2h_batch_sumsq: one pass, fixed order)."""
    x = x.contiguous()
    B = x.shape[0]
    v = torch.view_as_real(x) if x.is_complex() else x
    if v.dtype != torch.float64:
        raise TypeError("residuals are float64 / complex128")
    n = v.numel() // max(B, 1)
    out = torch.empty(B, dtype=torch.float64, device=x.device)
    lib = _lib.lib()
    work = torch.empty(max(int(lib.s2h_batch_sumsq_workspace_bytes(B)), 8), dtype=torch.uint8,
                       device=x.device)
    _lib.check(lib.s2h_batch_sumsq(v.data_ptr(), n, n, B, out.data_ptr(), work.data_ptr(),
                                   torch.cuda.current_stream(x.device).cuda_stream),
               "batch sumsq")
    return out


def _track(history, flm, flm_true, unpack=lambda x: x):
    if history is None or flm_true is None:
        return
    history.append(float((unpack(flm) - flm_true).abs().max() / flm_true.abs().max()))
