"""s2h_forward_noalloc (the forward with its workspace in a fixed device
array) against s2h_forward with a caller workspace, through the C ABI (needs a
GPU).  Two calls in a row on one device, the second with a smaller input, then
a workspace that fills the scratch exactly and one that exceeds it."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import signals
from s2fft_amd import _lib
from s2fft_amd.transforms import spherical as sph

pytestmark = pytest.mark.gpu
TOL = 1e-12  # FP64 transforms: max-relative difference between the entry points


def _say(msg):
    print(msg, flush=True)


def _maps(L, nside, sampling, spin, reality, seeds):
    flm = np.stack([signals.generate_flm(np.random.default_rng(s), L, spin=spin, reality=reality)
                    for s in seeds])
    plan = sph.get_plan(sampling, L, nside, spin, 0, reality, device=torch.device("cuda:0"))
    f = sph.inverse_hip(torch.from_numpy(flm).cuda(), L, spin, nside, sampling, reality,
                        plan=plan)
    if reality:
        f = f.real
    return plan, f.contiguous()


def _forward_with_work(plan, f, L):
    B = f.shape[0]
    flm = torch.empty((B, L, 2 * L - 1), dtype=torch.complex128, device=f.device)
    work = torch.empty(max(plan.workspace_bytes(B), 1), dtype=torch.uint8, device=f.device)
    _lib.check(_lib.lib().s2h_forward(plan.handle, f.data_ptr(), int(f.is_complex()),
                                      flm.data_ptr(), work.data_ptr(), B,
                                      torch.cuda.current_stream().cuda_stream), "forward")
    return flm


def _forward_noalloc(plan, f, L, out=None):
    B = f.shape[0]
    flm = out if out is not None else torch.empty((B, L, 2 * L - 1), dtype=torch.complex128,
                                                  device=f.device)
    st = _lib.lib().s2h_forward_noalloc(plan.handle, f.data_ptr(), int(f.is_complex()),
                                        flm.data_ptr(), B,
                                        torch.cuda.current_stream().cuda_stream)
    return st, flm


def _max_rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def test_scratch_size_reported():
    assert _lib.lib().s2h_noalloc_scratch_bytes() == 64 << 20


@pytest.mark.parametrize("sampling,L,nside,spin,reality,B1,B2", [
    ("healpix", 128, 64, 0, True, 3, 1),
    ("mw", 64, 0, 2, False, 2, 1),
    ("gl", 48, 0, 0, True, 4, 2),
])
def test_noalloc_matches_forward_two_calls(sampling, L, nside, spin, reality, B1, B2):
    _say(f"noalloc: {sampling} L={L} nside={nside} spin={spin} reality={reality} "
         f"B={B1} then {B2}")
    plan, f = _maps(L, nside, sampling, spin, reality, range(10, 10 + B1))
    assert plan.workspace_bytes(B1) <= _lib.lib().s2h_noalloc_scratch_bytes()
    ref1 = _forward_with_work(plan, f, L)
    st, got1 = _forward_noalloc(plan, f, L)
    _lib.check(st, "forward_noalloc")
    torch.cuda.synchronize()
    e1 = _max_rel(got1, ref1)
    # second call: a smaller input on the same device (the scratch keeps the
    # first call's contents)
    f2 = f[:B2].clone().contiguous() * 0.5
    ref2 = _forward_with_work(plan, f2, L)
    st, got2 = _forward_noalloc(plan, f2, L)
    _lib.check(st, "forward_noalloc")
    torch.cuda.synchronize()
    e2 = _max_rel(got2, ref2)
    _say(f"  call 1 max-rel {e1:.2e} (identical: {bool(torch.equal(got1, ref1))}), "
         f"call 2 max-rel {e2:.2e} (identical: {bool(torch.equal(got2, ref2))})")
    assert e1 <= TOL and e2 <= TOL


def test_noalloc_full_scratch_and_rejection():
    # nside 256, L 512, real: 16.76 MB of workspace per map, so 4 maps fill
    # 67.04 of the 67.11 MB scratch and 5 exceed it
    L, nside = 512, 256
    plan, f = _maps(L, nside, "healpix", 0, True, range(40, 45))
    cap = _lib.lib().s2h_noalloc_scratch_bytes()
    assert plan.workspace_bytes(4) <= cap < plan.workspace_bytes(5)
    _say("noalloc: nside 256 with 5 maps (rejected), then 4 maps (fills the scratch)")
    out = torch.zeros((5, L, 2 * L - 1), dtype=torch.complex128, device=f.device)
    st, _ = _forward_noalloc(plan, f, L, out=out)
    assert st == _lib.S2H_ERR_SCRATCH
    assert "scratch" in _lib.lib().s2h_status_string(st).decode()
    assert float(out.abs().max()) == 0.0  # nothing launched
    with pytest.raises(RuntimeError, match="scratch"):
        _lib.check(st, "forward_noalloc")
    f4 = f[:4].contiguous()
    ref = _forward_with_work(plan, f4, L)
    st, got = _forward_noalloc(plan, f4, L)
    _lib.check(st, "forward_noalloc")
    torch.cuda.synchronize()
    e = _max_rel(got, ref)
    _say(f"  4 maps max-rel {e:.2e}")
    assert e <= TOL


def test_noalloc_rejects_bad_arguments():
    lib = _lib.lib()
    assert lib.s2h_forward_noalloc(None, None, 0, None, 1, None) == _lib.S2H_ERR_INVALID
    plan, f = _maps(16, 0, "mw", 0, False, [1])
    flm = torch.empty((1, 16, 31), dtype=torch.complex128, device=f.device)
    assert lib.s2h_forward_noalloc(plan.handle, f.data_ptr(), 1, flm.data_ptr(), -1,
                                   None) == _lib.S2H_ERR_INVALID
    assert lib.s2h_forward_noalloc(plan.handle, f.data_ptr(), 1, flm.data_ptr(), 0,
                                   None) == _lib.S2H_OK
