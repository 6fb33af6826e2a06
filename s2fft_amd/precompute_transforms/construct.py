"""Kernel construction for the precompute path (s2fft/precompute_transforms/construct.py).

``spin_spherical_kernel`` and ``wigner_kernel`` keep the reference
signatures (construct.py:16-161 and 332-468) and return the reference's
kernels -- same layout, dtype and contents -- as torch tensors on the current
GPU, built by the HIP l-recursion (``s2h_precompute_kernel``):

* spin: K[t, l, m] = d^l_{m,-spin}(theta_t) sqrt((2l+1)/4pi), times the
  forward quadrature weight of ring t (forward kernels) and the HEALPix ring
  phase e^{-+ i m phi0(t)}; shape (n_theta, L, L or 2L-1), float64 or
  complex128 (HEALPix).  Forward MW/MWSS kernels live on the MWSS grid at 2L.
* Wigner: K[n, t, l, m] = d^l_{mn}(beta_t) times 2 pi / (2N-1) and the
  quadrature weight (forward) or (2l+1)/(8 pi^2) (inverse), and the HEALPix
  phase; shape (N or 2N-1, n_theta, L, 2L-1).

Like the reference's, the kernels are O(L^3) (O(N L^3)) in memory.
``recursion`` / ``mode`` select among the reference's CPU recursions; the
device recursion serves every choice, and the reference's argument errors
are reproduced.
"""

from warnings import warn

import numpy as np
import torch

from .. import _lib
from ..transforms import spherical as sph

PM_MAX_STABLE_SPIN = 6  # construct.py:13


def _build(plan, forward, scale, shape, cplx, device):
    out = torch.empty(shape, dtype=torch.complex128 if cplx else torch.float64, device=device)
    nbytes = int(_lib.lib().s2h_precompute_kernel_bytes(plan.handle, int(forward)))
    if nbytes != out.numel() * out.element_size():
        raise RuntimeError(f"kernel size mismatch ({nbytes} vs {out.numel() * out.element_size()})")
    _lib.check(_lib.lib().s2h_precompute_kernel(plan.handle, int(forward), float(scale),
                                                out.data_ptr(), sph._stream(device)),
               "precompute kernel")
    return out


def _ntheta(L, sampling, nside, forward):
    from ..sampling import s2_samples

    s = sampling.lower()
    if forward and s in ("mw", "mwss"):
        return 2 * L + 1
    return s2_samples.ntheta(L, s, nside)


def spin_spherical_kernel(L, spin=0, reality=False, sampling="mw", nside=None, forward=True,
                          recursion="auto", device=None):
    """Kernel of the spin-spherical precompute transforms (construct.py:16-161)."""
    if reality and spin != 0:
        reality = False
        warn("Reality acceleration only supports spin 0 fields. "
             "Defering to complex transform.", stacklevel=2)
    rec = recursion.lower()
    if rec == "price-mcewen" and abs(spin) > PM_MAX_STABLE_SPIN:
        raise ValueError("The Price-McEwen recursion can become unstable for spins >= "
                         f"{PM_MAX_STABLE_SPIN}.")
    if rec not in ("auto", "price-mcewen", "risbo"):
        raise ValueError(f"Recursion method {recursion} not recognised.")
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    s = sampling.lower()
    with torch.cuda.device(dev):
        plan = sph.get_plan(s, L, nside, spin, 0, reality, device=dev)
        m_dim = L if reality else 2 * L - 1
        shape = (_ntheta(L, s, nside, forward), L, m_dim)
        return _build(plan, forward, 1.0, shape, s == "healpix", dev)


def wigner_kernel(L, N, reality=False, sampling="mw", nside=None, forward=False, mode="auto",
                  device=None):
    """Kernel of the Wigner precompute transforms (construct.py:332-468)."""
    s = sampling.lower()
    if mode.lower() == "fft" and s not in ("mw", "mwss", "dh"):
        raise ValueError(f"Fourier based recursion is not valid for {sampling} sampling.")
    if mode.lower() not in ("auto", "direct", "fft"):
        raise ValueError(f"Recursion method {mode} not recognised.")
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    n0 = 0 if reality else -(N - 1)
    G = 2 * N - 1
    scale = 2 * np.pi / G if forward else 1.0
    planes = []
    with torch.cuda.device(dev):
        for n in range(n0, N):
            plan = sph.get_plan(s, L, nside, -n, 0, False, _lib.NORM_WIGNER, -1 if n % 2 else 1,
                                device=dev)
            shape = (_ntheta(L, s, nside, forward), L, 2 * L - 1)
            planes.append(_build(plan, forward, scale, shape, s == "healpix", dev))
        return torch.stack(planes)
