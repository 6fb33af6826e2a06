"""The XLA custom-call entry healpix_fft_hip through its own ABI (needs a GPU).

Replaces healpix_fft_cuda (lib/src/extensions.cc:26-100).  The descriptor is
made by s2h_build_healpix_fft_descriptor, which maps (forward, normalize) to a
norm as build_healpix_fft_descriptor does (extensions.cc:115-149) and prepares
the ring tables; the call then takes (stream, buffers, opaque, opaque_len).
Expected values: the oracle's ring FFTs (healpix_ffts.py:187-226, 335-371)
scaled per ring by the norm (1/n or 1/sqrt(n), lib/src/s2fft_callbacks.cu:
256-308).  Tolerances: 1e-12 for complex128 buffers; complex64 buffers hold
the result rounded to FP32 (the reference's CUDA tests use 1e-7 absolute and
relative, tests/test_healpix_ffts.py:70,86; here 1e-6 norm-wise).
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import ffts as offts
from tests.conftest import max_rel

from s2fft_amd import _lib

pytestmark = pytest.mark.gpu


def _call(desc, x, out):
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    bufs = (ctypes.c_void_p * 2)(x.data_ptr(), out.data_ptr())
    _lib.lib().healpix_fft_hip(stream, bufs, ctypes.string_at(ctypes.byref(desc),
                                                              ctypes.sizeof(desc)),
                               ctypes.sizeof(desc))
    return _lib.lib().healpix_fft_hip_last_status()


def _desc(nside, L, reality, forward, normalize, double):
    d = _lib.HealpixFftDescriptor()
    _lib.check(_lib.lib().s2h_build_healpix_fft_descriptor(nside, L, int(reality), int(forward),
                                                           int(normalize), int(double),
                                                           ctypes.byref(d)))
    return d


def _ring_scale(nside, L, mode):
    """Per-row factor of the (4 nside - 1, 2L) ftm: 1, 1/n or 1/sqrt(n)."""
    _, n = offts.ring_offsets(nside)
    n = n.astype(np.float64)
    return {0: np.ones_like(n), 1: 1.0 / n, 2: 1.0 / np.sqrt(n)}[mode]


@pytest.mark.parametrize("nside,L", [(4, 8), (5, 10), (8, 24)])
@pytest.mark.parametrize("normalize", [False, True])
@pytest.mark.parametrize("double", [True, False])
@pytest.mark.parametrize("reality", [False, True])
def test_healpix_fft_hip_forward_and_backward(rng, nside, L, normalize, double, reality):
    npix = 12 * nside * nside
    f = rng.standard_normal(npix) + (0 if reality else 1j * rng.standard_normal(npix))
    cdt = torch.complex128 if double else torch.complex64
    tol = 1e-12 if double else 1e-6
    # forward: normalize -> FORWARD norm (1/n per ring), else unscaled
    ref = offts.healpix_fft(f.astype(np.complex128), L, nside, reality)
    ref = ref * _ring_scale(nside, L, 1 if normalize else 0)[:, None]
    d = _desc(nside, L, reality, True, normalize, double)
    assert d.norm == (_lib.FFT_NORM_FORWARD if normalize else _lib.FFT_NORM_BACKWARD)
    x = torch.from_numpy(f).to(cdt).cuda()
    x_before = x.clone()
    ftm = torch.empty((4 * nside - 1, 2 * L), dtype=cdt, device="cuda")
    assert _call(d, x, ftm) == _lib.S2H_OK
    assert max_rel(ftm.cpu().numpy(), ref) < tol
    assert torch.equal(x, x_before)          # the input is never modified
    # backward: normalize -> BACKWARD norm (1/n per ring), else unscaled
    g = rng.standard_normal(ref.shape) + 1j * rng.standard_normal(ref.shape)
    back_ref = offts.healpix_ifft(g.copy(), L, nside, reality)
    offs, n = offts.ring_offsets(nside)
    if normalize:
        back_ref = back_ref / np.repeat(n, n)
    d = _desc(nside, L, reality, False, normalize, double)
    assert d.norm == (_lib.FFT_NORM_BACKWARD if normalize else _lib.FFT_NORM_FORWARD)
    y = torch.from_numpy(g).to(cdt).cuda()
    back = torch.empty(npix, dtype=cdt, device="cuda")
    assert _call(d, y, back) == _lib.S2H_OK
    assert max_rel(back.cpu().numpy(), back_ref) < tol


@pytest.mark.parametrize("norm", [3, 4])
def test_healpix_fft_ex_ortho_and_none(rng, norm):
    nside, L = 4, 8
    npix = 12 * nside * nside
    f = rng.standard_normal(npix) + 1j * rng.standard_normal(npix)
    mode = 2 if norm == _lib.FFT_NORM_ORTHO else 0
    lib = _lib.lib()
    s = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.s2h_healpix_fft_prepare(nside, L))
    x = torch.from_numpy(f).cuda()
    ftm = torch.empty((4 * nside - 1, 2 * L), dtype=torch.complex128, device="cuda")
    _lib.check(lib.s2h_healpix_fft_ex(nside, L, 0, norm, 1, x.data_ptr(), ftm.data_ptr(), 1, s))
    ref = offts.healpix_fft(f, L, nside) * _ring_scale(nside, L, mode)[:, None]
    assert max_rel(ftm.cpu().numpy(), ref) < 1e-12
    back = torch.empty(npix, dtype=torch.complex128, device="cuda")
    _lib.check(lib.s2h_healpix_ifft_ex(nside, L, 0, norm, 1, ftm.data_ptr(), back.data_ptr(), 1, s))
    # the inverse folds the 2L columns back onto each ring (healpix_ffts.py:19-82);
    # ORTHO scales each ring by 1/sqrt(n) again, NONE not at all
    _, n = offts.ring_offsets(nside)
    want = offts.healpix_ifft(ftm.cpu().numpy(), L, nside) * np.repeat(
        _ring_scale(nside, L, mode), n)
    assert max_rel(back.cpu().numpy(), want) < 1e-12


def test_healpix_fft_hip_rejects_bad_descriptors():
    nside, L = 4, 8
    x = torch.zeros(12 * nside * nside, dtype=torch.complex128, device="cuda")
    out = torch.zeros((4 * nside - 1, 2 * L), dtype=torch.complex128, device="cuda")
    d = _desc(nside, L, False, True, False, True)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    bufs = (ctypes.c_void_p * 2)(x.data_ptr(), out.data_ptr())
    raw = ctypes.string_at(ctypes.byref(d), ctypes.sizeof(d))
    _lib.lib().healpix_fft_hip(stream, bufs, raw, len(raw) - 4)     # wrong length
    assert _lib.lib().healpix_fft_hip_last_status() == _lib.S2H_ERR_INVALID
    d.shift = 0
    assert _call(d, x, out) == _lib.S2H_ERR_UNSUPPORTED
    d.shift, d.norm = 1, 7
    assert _call(d, x, out) == _lib.S2H_ERR_INVALID
    # tables never prepared for this (nside, L): the launch refuses instead of allocating
    e = _lib.HealpixFftDescriptor(6, 12, 0, 1, _lib.FFT_NORM_BACKWARD, 1, 1, 1)
    xe = torch.zeros(12 * 36, dtype=torch.complex128, device="cuda")
    oe = torch.zeros((23, 24), dtype=torch.complex128, device="cuda")
    assert _call(e, xe, oe) == _lib.S2H_ERR_INVALID
