"""ctypes binding of ``libs2fft_amd.so`` (the C ABI declared in include/s2fft_amd.h).

The library is built in-tree by ``python -m s2fft_amd.build``.  There is no
fallback: if it is missing every transform raises ``RuntimeError``.
"""

import ctypes
import threading
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libs2fft_amd.so"
_lock = threading.Lock()
_lib = None

S2H_MW, S2H_MWSS, S2H_DH, S2H_GL, S2H_HEALPIX = range(5)
SAMPLING_CODES = {"mw": S2H_MW, "mwss": S2H_MWSS, "dh": S2H_DH, "gl": S2H_GL,
                  "healpix": S2H_HEALPIX}
NORM_SPIN, NORM_WIGNER, NORM_NONE = 0, 1, 2

S2H_OK = 0
S2H_ERR_INVALID = -1
S2H_ERR_UNSUPPORTED = -2
S2H_ERR_HIP = -3
S2H_ERR_NOMEM = -4
S2H_ERR_SPIN = -5
S2H_ERR_BANDLIMIT = -6
S2H_ERR_SCRATCH = -7


class Desc(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int32) for name in
                ("sampling", "L", "nside", "spin", "L_lower", "reality", "norm", "extra_sign")]


class HealpixFftDescriptor(ctypes.Structure):
    """s2h_healpix_fft_descriptor: the reference's s2fftDescriptor fields as int32."""
    _fields_ = [(name, ctypes.c_int32) for name in
                ("nside", "L", "reality", "forward", "norm", "shift", "double_precision",
                 "batch")]


FFT_NORM_FORWARD, FFT_NORM_BACKWARD, FFT_NORM_ORTHO, FFT_NORM_NONE = 1, 2, 3, 4


EXPORTED = (
    "s2h_status_string", "s2h_version", "s2h_plan_create", "s2h_plan_destroy",
    "s2h_workspace_bytes", "s2h_inverse_output_is_complex", "s2h_forward", "s2h_inverse",
    "s2h_lat_forward", "s2h_lat_inverse", "s2h_healpix_fft", "s2h_healpix_ifft",
    "healpix_fft_hip", "healpix_fft_hip_last_status", "s2h_gamma_dft",
    "s2h_healpix_fft_prepare", "s2h_healpix_fft_ex", "s2h_healpix_ifft_ex",
    "s2h_build_healpix_fft_descriptor",
    "s2h_wigner_plan_create", "s2h_wigner_plan_destroy", "s2h_wigner_workspace_bytes",
    "s2h_wigner_forward", "s2h_wigner_inverse", "s2h_wigner_forward_adjoint",
    "s2h_wigner_inverse_adjoint", "s2h_compute_all_slices", "s2h_precompute_kernel_bytes",
    "s2h_precompute_kernel", "s2h_precompute_forward", "s2h_precompute_inverse",
    "s2h_profile_enable", "s2h_profile_query", "s2h_wigner_reality_fill",
    "s2h_batch_sumsq_workspace_bytes", "s2h_batch_sumsq",
    "s2h_ring_spectra_bytes", "s2h_ring_analysis", "s2h_forward_rings", "s2h_inverse_rings",
    "s2h_ring_residual", "s2h_flm_pack", "s2h_flm_unpack",
    "s2h_ring_residual_scaled", "s2h_ring_dot_workspace_bytes", "s2h_ring_dot",
    "s2h_ring_residual_ex", "s2h_ring_dot_ex", "s2h_inverse_rings_dot_workspace_bytes",
    "s2h_inverse_rings_dot",
    "s2h_flm_dots_workspace_bytes", "s2h_flm_dots", "s2h_cg_update", "s2h_cg_finish",
    "s2h_forward_adjoint", "s2h_inverse_adjoint", "s2h_probe_mfma_f64", "s2h_probe_mfma_f64_rate",
    "s2h_probe_mfma_f32", "s2h_forward_ex", "s2h_inverse_ex",
    "s2h_mlp_param_count", "s2h_mlp_workspace_bytes", "s2h_mlp_precondition",
    "s2h_mlp_precondition_dot",
    "s2h_mlp_features",
    "s2h_mlp_grad_workspace_bytes", "s2h_mlp_grad", "s2h_mlp_adam", "s2h_probe_mfma_bf16",
    "s2h_noalloc_scratch_bytes", "s2h_forward_noalloc",
)

KERNEL_LAT_FORWARD, KERNEL_LAT_INVERSE, KERNEL_RING_FFT_FORWARD = 0, 1, 2
KERNEL_RING_FFT_INVERSE, KERNEL_RESAMPLE = 3, 4
KERNEL_LAT_FORWARD_FP32, KERNEL_LAT_INVERSE_FP32 = 5, 6
KERNEL_LAT_FORWARD_BF16, KERNEL_LAT_INVERSE_BF16 = 7, 8
KERNEL_RING_RESIDUAL = 9
PREC_FP64, PREC_MIXED, PREC_BF16, PREC_BF16X1 = 0, 1, 2, 3
FWD_ACCUMULATE = 1
FLM_PACKED = 4
SPECTRA_TILED = 8
SPECTRA_G_RING_MAJOR = 16
INV_REAL_OUTPUT, INV_SUBTRACT = 1, 2


def _declare(lib):
    vp, i32, i64, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_size_t
    sig = {
        "s2h_status_string": (ctypes.c_char_p, [i32]),
        "s2h_version": (i32, []),
        "s2h_plan_create": (i32, [ctypes.POINTER(vp), ctypes.POINTER(Desc)]),
        "s2h_plan_destroy": (i32, [vp]),
        "s2h_workspace_bytes": (sz, [vp, i32]),
        "s2h_inverse_output_is_complex": (i32, [vp]),
        "s2h_forward": (i32, [vp, vp, i32, vp, vp, i32, vp]),
        "s2h_inverse": (i32, [vp, vp, vp, vp, i32, vp]),
        "s2h_forward_ex": (i32, [vp, vp, i32, vp, vp, i32, i32, i32, vp]),
        "s2h_inverse_ex": (i32, [vp, vp, vp, vp, i32, i32, i32, vp, vp]),
        "s2h_forward_adjoint": (i32, [vp, vp, vp, vp, i32, vp]),
        "s2h_probe_mfma_f64": (i32, [vp, vp, vp, vp]),
        "s2h_probe_mfma_f64_rate": (i32, [i32, i32, vp, vp]),
        "s2h_probe_mfma_f32": (i32, [vp, vp, vp, vp]),
        "s2h_probe_mfma_bf16": (i32, [vp, vp, vp, vp]),
        "s2h_mlp_param_count": (i32, []),
        "s2h_mlp_workspace_bytes": (ctypes.c_size_t, [i32]),
        "s2h_mlp_precondition": (i32, [vp, vp, vp, i32, i32, i32, i32, vp, vp]),
        "s2h_mlp_precondition_dot": (i32, [vp, vp, vp, i32, i32, i32, i32, vp, vp, vp]),
        "s2h_mlp_features": (i32, [vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]),
        "s2h_mlp_grad_workspace_bytes": (sz, [i32]),
        "s2h_mlp_grad": (i32, [vp, vp, vp, vp, i64, ctypes.c_float, vp, vp, vp, i32, vp]),
        "s2h_mlp_adam": (i32, [vp, vp, vp, vp, i32, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, ctypes.c_float, vp]),
        "s2h_inverse_adjoint": (i32, [vp, vp, i32, vp, vp, i32, vp]),
        "s2h_lat_forward": (i32, [vp, vp, vp, i32, vp]),
        "s2h_lat_inverse": (i32, [vp, vp, vp, i32, vp]),
        "s2h_healpix_fft": (i32, [i32, i32, i32, vp, vp, i32, vp]),
        "s2h_healpix_ifft": (i32, [i32, i32, i32, vp, vp, i32, vp]),
        "s2h_healpix_fft_prepare": (i32, [i32, i32]),
        "s2h_healpix_fft_ex": (i32, [i32, i32, i32, i32, i32, vp, vp, i32, vp]),
        "s2h_healpix_ifft_ex": (i32, [i32, i32, i32, i32, i32, vp, vp, i32, vp]),
        "s2h_build_healpix_fft_descriptor": (i32, [i32, i32, i32, i32, i32, i32,
                                                   ctypes.POINTER(HealpixFftDescriptor)]),
        "healpix_fft_hip": (None, [vp, ctypes.POINTER(vp), ctypes.c_char_p, sz]),
        "s2h_compute_all_slices": (i32, [vp, vp, i32, vp, vp]),
        "s2h_precompute_kernel_bytes": (sz, [vp, i32]),
        "s2h_precompute_kernel": (i32, [vp, i32, ctypes.c_double, vp, vp]),
        "s2h_precompute_forward": (i32, [vp, vp, i32, ctypes.c_double, vp, i32, vp, vp, i32,
                                         vp]),
        "s2h_precompute_inverse": (i32, [vp, vp, i32, ctypes.c_double, vp, vp, vp, i32, vp]),
        "s2h_wigner_plan_create": (i32, [ctypes.POINTER(vp), vp]),
        "s2h_wigner_plan_destroy": (i32, [vp]),
        "s2h_wigner_workspace_bytes": (sz, [vp]),
        "s2h_wigner_forward": (i32, [vp, vp, i32, vp, vp, vp]),
        "s2h_wigner_inverse": (i32, [vp, vp, vp, vp, vp]),
        "s2h_wigner_forward_adjoint": (i32, [vp, vp, vp, vp, vp]),
        "s2h_wigner_inverse_adjoint": (i32, [vp, vp, i32, vp, vp, vp]),
        "healpix_fft_hip_last_status": (i32, []),
        "s2h_gamma_dft": (i32, [i32, i64, i32, ctypes.c_double, vp, vp, vp]),
        "s2h_wigner_reality_fill": (i32, [vp, i32, i32, vp]),
        "s2h_batch_sumsq_workspace_bytes": (ctypes.c_size_t, [i32]),
        "s2h_batch_sumsq": (i32, [vp, i64, i64, i32, vp, vp, vp]),
        "s2h_ring_spectra_bytes": (sz, [vp, i32]),
        "s2h_ring_analysis": (i32, [vp, vp, i32, vp, i32, vp]),
        "s2h_forward_rings": (i32, [vp, vp, vp, i32, i32, i32, vp]),
        "s2h_inverse_rings": (i32, [vp, vp, vp, i32, i32, i32, vp]),
        "s2h_flm_pack": (i32, [i32, vp, vp, i32, vp]),
        "s2h_flm_unpack": (i32, [i32, vp, vp, i32, vp]),
        "s2h_ring_residual": (i32, [vp, vp, vp, vp, i32, vp]),
        "s2h_ring_residual_scaled": (i32, [vp, vp, vp, vp, vp, vp, vp, i32, vp]),
        "s2h_ring_dot_workspace_bytes": (sz, [vp, i32]),
        "s2h_ring_dot": (i32, [vp, vp, vp, vp, vp, vp, i32, vp]),
        "s2h_ring_residual_ex": (i32, [vp, vp, vp, vp, vp, vp, vp, i32, i32, vp]),
        "s2h_ring_dot_ex": (i32, [vp, vp, vp, vp, vp, vp, i32, i32, vp]),
        "s2h_inverse_rings_dot_workspace_bytes": (sz, [vp, i32]),
        "s2h_inverse_rings_dot": (i32, [vp, vp, vp, i32, i32, i32, vp, vp, vp, vp, vp]),
        "s2h_flm_dots_workspace_bytes": (sz, [i32, i32]),
        "s2h_flm_dots": (i32, [vp, ctypes.POINTER(vp), i32, i64, i32, i32, vp, vp, vp]),
        "s2h_cg_update": (i32, [vp, vp, vp, vp, vp, vp, vp, vp, i64, i32, vp]),
        "s2h_cg_finish": (i32, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), i32, i64, i32, vp,
                                i32, vp]),
        "s2h_noalloc_scratch_bytes": (sz, []),
        "s2h_forward_noalloc": (i32, [vp, vp, i32, vp, i32, vp]),
        "s2h_profile_enable": (i32, [i32]),
        "s2h_profile_query": (i32, [i32, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Load the shared library once (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not _LIB_PATH.exists():
                    raise RuntimeError(
                        f"s2fft_amd native library not found at {_LIB_PATH}; build it with "
                        "`python -m s2fft_amd.build` (there is no CPU fallback)")
                handle = ctypes.CDLL(str(_LIB_PATH))
                _declare(handle)
                _lib = handle
    return _lib


def library_path():
    return _LIB_PATH


def check(status, what="s2fft_amd"):
    """Map an s2h status to the reference's Python exception types."""
    if status == S2H_OK:
        return
    msg = lib().s2h_status_string(status).decode()
    if status == S2H_ERR_SPIN:
        raise Warning(msg)
    if status == S2H_ERR_BANDLIMIT:
        raise AssertionError(msg)
    if status == S2H_ERR_INVALID:
        raise ValueError(f"{what}: {msg}")
    if status == S2H_ERR_UNSUPPORTED:
        raise NotImplementedError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg} (status {status})")


class Plan:
    """Owner of one native plan (constant tables on the current device)."""

    def __init__(self, sampling, L, nside=0, spin=0, L_lower=0, reality=False,
                 norm=NORM_SPIN, extra_sign=1):
        d = Desc(SAMPLING_CODES[sampling], L, nside or 0, spin, L_lower, int(bool(reality)),
                 norm, extra_sign)
        handle = ctypes.c_void_p()
        check(lib().s2h_plan_create(ctypes.byref(handle), ctypes.byref(d)), "plan")
        self.handle = handle
        self.desc = d

    def workspace_bytes(self, batch):
        return int(lib().s2h_workspace_bytes(self.handle, batch))

    def inverse_output_is_complex(self):
        return bool(lib().s2h_inverse_output_is_complex(self.handle))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib is not None:
            _lib.s2h_plan_destroy(h)
            self.handle = None


def profile_enable(on=True):
    check(lib().s2h_profile_enable(int(bool(on))), "profile")


def profile_query(kernel):
    """(total milliseconds, launch count) of one kernel kind since enabling."""
    ms = ctypes.c_double()
    n = ctypes.c_int()
    check(lib().s2h_profile_query(kernel, ctypes.byref(ms), ctypes.byref(n)), "profile")
    return ms.value, n.value
