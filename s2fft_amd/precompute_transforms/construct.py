"""Kernel construction for the precompute API (s2fft/precompute_transforms/construct.py).

``spin_spherical_kernel`` / ``wigner_kernel`` keep the reference signatures
(construct.py:16-24 and 332-341) but return a :class:`KernelPlan`: the
device-resident plan (geometry, recursion coefficients, ring tables) that the
transforms use.  Its memory is O(L^2) where the reference kernel is O(L^3)
(construct.py:28-30).
"""

from dataclasses import dataclass

from ..transforms import spherical as sph


@dataclass(frozen=True)
class KernelPlan:
    """Stands in for the reference's d-kernel array."""

    L: int
    spin: int
    reality: bool
    sampling: str
    nside: object
    forward: bool
    N: int = 0          # > 0 for a Wigner kernel

    def plan(self, device="cuda"):
        return sph.get_plan(self.sampling.lower(), self.L, self.nside, self.spin, 0,
                            self.reality, device=device)


def spin_spherical_kernel(L, spin=0, reality=False, sampling="mw", nside=None, forward=True,
                          recursion="auto"):
    """Plan handle for the spin-spherical transforms (construct.py:16-161).

    ``recursion`` is accepted for signature compatibility: the d-values always
    come from the device l-recursion."""
    if reality and spin != 0:
        reality = False
    return KernelPlan(L, spin, bool(reality), sampling, nside, bool(forward))


def wigner_kernel(L, N, reality=False, sampling="mw", nside=None, forward=False, mode="auto"):
    """Plan handle for the Wigner transforms (construct.py:332-468)."""
    return KernelPlan(L, 0, bool(reality), sampling, nside, bool(forward), N)
