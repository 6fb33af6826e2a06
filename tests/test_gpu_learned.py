"""The learned preconditioner end to end (needs a GPU).

Trains the network briefly on the GPU (plain-transform data at nside 32) and
compares the mixed refinement with and without it on held-out maps at
nside 64: the Jacobi-judged learned steps must reach 1e-7 in no more steps
than Jacobi, never raise a map's residual, and the counts are printed.
"""

import pytest
import torch

from s2fft_amd.learned import precond as lp
from s2fft_amd.learned import train_healpix as th

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_model():
    X, Y = lp.make_training_rows(32, 128, seed=77, keep=0.5)
    model, hist = lp.train(X, Y, steps=400, rows_per_step=16384, lr=3e-3, seed=1)
    return model, hist


def test_training_loss_falls(small_model):
    _, hist = small_model
    assert hist[-1][1] < 0.5 * hist[0][1]


@pytest.mark.parametrize("precision", ["bf16", "mixed"])
def test_learned_gain_over_jacobi(small_model, precision):
    model, _ = small_model
    out = th.eval_gain(model, 64, n_maps=4, precision=precision, max_iter=14)
    jac, lrn = out["jacobi"], out["learned"]
    print(precision, "jacobi steps", jac["steps_to_tol"], "learned steps", lrn["steps_to_tol"],
          "accepted", lrn["stats"])
    print("jacobi errs ", ["%.1e" % e for e in jac["errs"]])
    print("learned errs", ["%.1e" % e for e in lrn["errs"]])
    assert jac["steps_to_tol"] is not None and lrn["steps_to_tol"] is not None
    assert lrn["steps_to_tol"] <= jac["steps_to_tol"]
