"""Writes healpix_nside512_L1024_iter3.npz: the oracle's HEALPix transform chain
of the nside-512 real spin-0 workload (seed 1), sampled.

    flm  = generate_flm(default_rng(1), L=1024, reality=True)
    f    = Re(inverse(flm))                       (HEALPix nside 512)
    flmk = forward with k steps of iterative refinement, k = 0..3

Stored: the max-relative error of flmk against flm for every k (the error the
reference's own refinement reaches at iter = k), the refined coefficients
flm3 at 4096 fixed (l, m >= 0) positions, the map f at 4096 fixed pixels and
max|f|, max|flm|.  The oracle restates s2fft/transforms/spherical.py:112-208,
467-588 and s2fft/utils/iterative_refinement.py:9-46 (see oracle/).

Run from the repository root: python -m tests.golden.make_healpix_nside512_iter3
(about 20 minutes on one core).
"""

import time
from pathlib import Path

import numpy as np

from oracle import signals
from oracle import transforms as otr

NSIDE, L, SEED, N_ITER = 512, 1024, 1, 3
OUT = Path(__file__).resolve().parent / "healpix_nside512_L1024_iter3.npz"


def main():
    t0 = time.time()
    flm = signals.generate_flm(np.random.default_rng(SEED), L, reality=True)
    f = np.real(otr.inverse(flm, L, 0, NSIDE, "healpix", True))
    print(f"inverse done ({time.time() - t0:.0f} s)", flush=True)
    den = np.max(np.abs(flm))
    fwd = lambda g: otr.forward(g, L, 0, NSIDE, "healpix", True)
    est = fwd(f)
    errs = [float(np.max(np.abs(est - flm)) / den)]
    print(f"iter 0: {errs[-1]:.3e} ({time.time() - t0:.0f} s)", flush=True)
    for k in range(N_ITER):
        est = est + fwd(f - np.real(otr.inverse(est, L, 0, NSIDE, "healpix", True)))
        errs.append(float(np.max(np.abs(est - flm)) / den))
        print(f"iter {k + 1}: {errs[-1]:.3e} ({time.time() - t0:.0f} s)", flush=True)

    pick = np.random.default_rng(12345)
    ell = pick.integers(0, L, 4096)
    m = (pick.random(4096) * (ell + 1)).astype(np.int64)       # 0 <= m <= ell
    pix = pick.choice(f.size, 4096, replace=False)
    np.savez(OUT, nside=NSIDE, L=L, seed=SEED, n_iter=N_ITER, err_by_iter=np.array(errs),
             ell=ell, m=m, flm_iter3=est[ell, L - 1 + m], pix=pix, f_pix=f[pix],
             f_max=np.max(np.abs(f)), flm_max=den, flm_iter3_max=np.max(np.abs(est)))
    print(f"wrote {OUT} ({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
