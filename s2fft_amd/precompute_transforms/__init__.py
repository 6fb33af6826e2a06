"""Drop-in for s2fft.precompute_transforms with ``method="hip"``.

The reference's precompute path (s2fft/precompute_transforms/) stores a full
Wigner-d kernel of O(L^3) memory and applies it as an einsum.  Here the same
transforms run on the batched matrix-core latitudinal kernels, which
regenerate the d-values on the device: a batch of maps turns each m's
projection into a dense contraction without a stored kernel.  ``construct``
returns a plan handle in place of the kernel array.
"""

from . import construct, spherical, wigner  # noqa: F401
