"""C ABI and host-side checks that need no GPU."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import s2fft_amd
from s2fft_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "s2fft_amd.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    keywords = {"if", "sizeof", "defined", "extern"}
    return sorted({n for n in names if n not in keywords and not n.isupper()})


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_functions()
    assert "s2h_forward" in names and "healpix_fft_hip" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.EXPORTED) == set(names)


def test_status_strings():
    lib = _lib.lib()
    for code in range(-7, 1):
        assert lib.s2h_status_string(code).decode() != "unknown status"
    assert lib.s2h_status_string(-99).decode() == "unknown status"
    assert lib.s2h_version() >= 1


def test_null_arguments_rejected_without_gpu():
    lib = _lib.lib()
    assert lib.s2h_plan_create(None, None) == _lib.S2H_ERR_INVALID
    assert lib.s2h_forward(None, None, 0, None, None, 1, None) == _lib.S2H_ERR_INVALID
    assert lib.s2h_inverse(None, None, None, None, 1, None) == _lib.S2H_ERR_INVALID
    assert lib.s2h_healpix_fft(0, 8, 0, None, None, 1, None) == _lib.S2H_ERR_INVALID
    assert lib.s2h_gamma_dft(9, 4, 1, 1.0, None, None, None) == _lib.S2H_ERR_INVALID
    assert lib.s2h_workspace_bytes(None, 1) == 0


def test_status_to_exception_mapping():
    with pytest.raises(Warning):
        _lib.check(_lib.S2H_ERR_SPIN)
    with pytest.raises(AssertionError):
        _lib.check(_lib.S2H_ERR_BANDLIMIT)
    with pytest.raises(ValueError):
        _lib.check(_lib.S2H_ERR_INVALID)
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.S2H_ERR_UNSUPPORTED)
    with pytest.raises(RuntimeError):
        _lib.check(_lib.S2H_ERR_HIP)


def test_api_argument_errors_before_any_device_work():
    """Reference error conventions: spherical.py:85-89, 425-429."""
    flm = np.zeros((4, 7), dtype=np.complex128)
    with pytest.raises(ValueError, match="not recognised"):
        s2fft_amd.inverse(flm, 4, method="numpy")
    with pytest.raises(ValueError, match="not recognised"):
        s2fft_amd.forward(np.zeros((4, 7)), 4, method="jax")
    with pytest.raises(Warning):
        s2fft_amd.inverse(np.zeros((10, 19), dtype=np.complex128), 10, spin=8, method="hip")
    with pytest.raises(Warning):
        s2fft_amd.forward(np.zeros((10, 19)), 10, spin=9, method="hip")
    with pytest.raises(ValueError):
        s2fft_amd.forward(np.zeros(48), 4, sampling="healpix", method="hip")


def test_precompute_api_errors_before_any_device_work():
    """The reference's argument errors (precompute_transforms/spherical.py:62-63,
    construct.py:65-69, 372-375) fire before anything touches a GPU."""
    from s2fft_amd.precompute_transforms import construct, spherical as pspin, wigner as pwig
    with pytest.raises(ValueError, match="not recognised"):
        pspin.forward(np.zeros((8, 15)), 8, method="numpy")
    with pytest.raises(ValueError, match="not recognised"):
        pwig.inverse(np.zeros((3, 8, 15)), 8, 2, method="jax")
    with pytest.raises(ValueError, match="unstable"):
        construct.spin_spherical_kernel(8, 7, recursion="price-mcewen")
    with pytest.raises(ValueError, match="not recognised"):
        construct.spin_spherical_kernel(8, 0, recursion="bogus")
    with pytest.raises(ValueError, match="not valid"):
        construct.wigner_kernel(8, 2, sampling="gl", mode="fft")


def test_cg_entry_points_reject_bad_arguments_without_gpu():
    lib = _lib.lib()
    inv = _lib.S2H_ERR_INVALID
    assert lib.s2h_ring_residual_scaled(None, None, None, None, None, None, None, 1, None) == inv
    assert lib.s2h_ring_dot(None, None, None, None, None, None, 1, None) == inv
    assert lib.s2h_ring_dot_workspace_bytes(None, 4) == 0
    assert lib.s2h_flm_dots_workspace_bytes(0, 4) == 0
    assert lib.s2h_flm_dots_workspace_bytes(5, 4) == 0
    assert lib.s2h_flm_dots_workspace_bytes(2, 0) == 0
    assert lib.s2h_flm_dots_workspace_bytes(2, 3) > 0
    vecs = (ctypes.c_void_p * 1)(None)
    assert lib.s2h_flm_dots(None, vecs, 0, 10, 0, 1, None, None, None) == inv  # nv < 1
    assert lib.s2h_flm_dots(None, vecs, 5, 10, 0, 1, None, None, None) == inv  # nv > 4
    assert lib.s2h_flm_dots(None, vecs, 1, 10, 0, 1, None, None, None) == inv  # null z / out
    assert lib.s2h_flm_dots(None, vecs, 1, 10, 0, 0, None, None, None) == _lib.S2H_OK  # no maps
    assert lib.s2h_cg_update(None, None, None, None, None, None, None, None, 10, 1, None) == inv
    assert lib.s2h_cg_update(None, None, None, None, None, None, None, None, 0, 1, None) == _lib.S2H_OK


def test_mlp_and_tiled_entry_points_reject_bad_arguments_without_gpu():
    lib = _lib.lib()
    inv = _lib.S2H_ERR_INVALID
    # RMS partial sums (1024 per map) and <delta, h> partial sums (256 per map)
    assert lib.s2h_mlp_workspace_bytes(0) == 0
    assert lib.s2h_mlp_workspace_bytes(3) == 3 * (1024 + 256) * 8
    buf = ctypes.create_string_buffer(64)
    ptr = ctypes.cast(buf, ctypes.c_void_p)
    # a null dot output, a null input, h == delta, a bad layout code
    assert lib.s2h_mlp_precondition_dot(ptr, ptr, ptr, 8, 2, 0, 1, None, ptr, None) == inv
    assert lib.s2h_mlp_precondition_dot(ptr, None, ptr, 8, 2, 0, 1, ptr, ptr, None) == inv
    assert lib.s2h_mlp_precondition(ptr, ptr, ptr, 8, 2, 0, 1, ptr, None) == inv
    assert lib.s2h_mlp_precondition(ptr, ptr, ptr, 8, 3, 0, 1, ptr, None) == inv
    assert lib.s2h_mlp_precondition(ptr, ptr, ptr, 8, 2, 8, 1, ptr, None) == inv  # lmin >= L
    # no maps: nothing to do, no device work
    assert lib.s2h_mlp_precondition_dot(ptr, ptr, ptr, 8, 2, 0, 0, None, None, None) == _lib.S2H_OK
    # the tiled ring-spectra entry points need a plan
    assert lib.s2h_ring_residual_ex(None, None, None, None, None, None, None, 1,
                                    _lib.SPECTRA_TILED, None) == inv
    assert lib.s2h_ring_dot_ex(None, None, None, None, None, None, 1, _lib.SPECTRA_TILED,
                               None) == inv
    assert lib.s2h_inverse_rings_dot(None, None, None, 1, 0, _lib.SPECTRA_TILED, None, None,
                                     None, None, None) == inv
    assert lib.s2h_inverse_rings_dot_workspace_bytes(None, 4) == 0


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], cwd=root,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()
