"""Matrix-core latitudinal kernels at a size with several ring groups per m.

HEALPix nside 256 has 512 ring pairs, so the syntheses split every m into two
workgroups of 256 pairs, dispatched as consecutive blocks of one XCD.  L = 515
gives 515 m-values, 3 past the last multiple of 8, where the block-to-(m,
group) map takes its tail branch.  The batched FP64 kernels (4 maps) must
agree with the single-map vector-ALU kernels to 1e-12 of max |value|: they
run the same recursion with different summation orders.  The BF16 syntheses
must stay within the tolerances stated in test_gpu_mixed.py.
"""

import numpy as np
import pytest
import torch

from oracle import signals
from tests.conftest import max_rel

from s2fft_amd.transforms import spherical as sph

pytestmark = pytest.mark.gpu

NSIDE, L, B = 256, 515, 4


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(515)
    flm = np.stack([signals.generate_flm(rng, L, 0, 0, True) for _ in range(B)])
    flm_t = torch.from_numpy(flm).cuda()
    plan = sph.get_plan("healpix", L, NSIDE, 0, 0, True, device=flm_t.device)
    return flm_t, plan


def test_batched_fp64_synthesis_matches_single_map(data):
    flm, plan = data
    Y = sph.inverse_rings_hip(flm, plan)
    for b in range(B):
        y1 = sph.inverse_rings_hip(flm[b:b + 1], plan)
        err = max_rel(Y[b:b + 1].cpu().numpy(), y1.cpu().numpy())
        assert err < 1e-12, (b, err)


@pytest.mark.parametrize("precision,tol", [("bf16", 1e-4), ("bf16x1", 2e-2), ("mixed", 2e-6)])
def test_reduced_synthesis_with_ring_groups(data, precision, tol):
    flm, plan = data
    Y64 = sph.inverse_rings_hip(flm, plan)
    Y = sph.inverse_rings_hip(flm, plan, precision)
    err = max_rel(Y.cpu().numpy(), Y64.cpu().numpy())
    assert 1e-12 < err < tol, err


def test_batched_fp64_projection_matches_single_map(data):
    flm, plan = data
    f = sph.inverse_hip(flm, L, 0, NSIDE, "healpix", True, plan=plan).real.contiguous()
    G = sph.ring_analysis_hip(f, plan)
    x = sph.forward_rings_hip(G, L, plan)
    for b in range(B):
        x1 = sph.forward_rings_hip(G[b:b + 1], L, plan)
        err = max_rel(x[b:b + 1].cpu().numpy(), x1.cpu().numpy())
        assert err < 1e-12, (b, err)
