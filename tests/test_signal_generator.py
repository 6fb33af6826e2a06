"""The product's seeded generators draw exactly what the oracle's do (CPU).

The oracle's draw order is pinned against raw generator calls in
tests/test_oracle_pins.py (test_generate_flm_draw_order_and_symmetry,
test_generate_flmn_draw_order_and_symmetry)."""

import numpy as np
import pytest

from oracle import signals
from s2fft_amd.utils import signal_generator as sg


@pytest.mark.parametrize("L", [1, 2, 7, 64])
@pytest.mark.parametrize("spin", [0, 2])
@pytest.mark.parametrize("L_lower", [0, 3])
@pytest.mark.parametrize("reality", [False, True])
def test_generate_flm_matches_oracle(L, spin, L_lower, reality):
    if max(spin, L_lower) > L:
        # no coefficients exist: both raise as the reference's draw would
        for gen in (sg.generate_flm, signals.generate_flm):
            with pytest.raises(ValueError):
                gen(np.random.default_rng(11), L, L_lower, spin, reality)
        return
    a = sg.generate_flm(np.random.default_rng(11), L, L_lower, spin, reality)
    b = signals.generate_flm(np.random.default_rng(11), L, L_lower, spin, reality)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("L,N", [(2, 1), (6, 3), (64, 4)])
@pytest.mark.parametrize("L_lower", [0, 2])
@pytest.mark.parametrize("reality", [False, True])
def test_generate_flmn_matches_oracle(L, N, L_lower, reality):
    a = sg.generate_flmn(np.random.default_rng(12), L, N, L_lower, reality)
    b = signals.generate_flmn(np.random.default_rng(12), L, N, L_lower, reality)
    assert np.array_equal(a, b)
    # the generator state after the call matches too (same number of draws)
    ra, rb = np.random.default_rng(12), np.random.default_rng(12)
    sg.generate_flmn(ra, L, N, L_lower, reality)
    signals.generate_flmn(rb, L, N, L_lower, reality)
    assert ra.standard_normal() == rb.standard_normal()
