"""Wigner-d tables on the device (s2fft/recursions/price_mcewen.py:377-643).

``compute_all_slices`` returns the same table as the reference's, dl[L-1+m,
t, l] = d^l_{m,-spin}(beta_t) for every m, sampled beta and l (zero for
l < max(|m|, |spin|)), computed by the HIP l-recursion
(``s2h_compute_all_slices``) instead of the reference's m-recursion.  The
table is O(L^3): meant for small band-limits, as the reference says.
"""

import numpy as np
import torch

from .. import _lib
from ..transforms import spherical as sph


def compute_all_slices(beta, L, spin=0, precomps=None):
    """dl (2L-1, len(beta), L) float64 (s2fft/recursions/price_mcewen.py:377-643).

    ``beta``: NumPy array (the result is NumPy) or torch tensor (the result
    stays on its GPU, or returns to the CPU for a CPU tensor).  ``precomps``
    is accepted for signature compatibility: the recursion tables live in a
    cached plan.
    """
    b, restore = sph._as_device_tensor(beta)
    b = b.reshape(-1).to(torch.float64).contiguous()
    with torch.cuda.device(b.device):
        plan = sph.get_plan("mw", L, None, spin, 0, False, _lib.NORM_NONE, device=b.device)
        dl = torch.empty((2 * L - 1, b.numel(), L), dtype=torch.float64, device=b.device)
        _lib.check(_lib.lib().s2h_compute_all_slices(plan.handle, b.data_ptr(), b.numel(),
                                                     dl.data_ptr(), sph._stream(b.device)),
                   "compute_all_slices")
    return restore(dl)
