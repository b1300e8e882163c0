"""ctypes binding of ``libteamtomo_amd.so`` (the C ABI in include/teamtomo_amd.h).

The library is required: importing the operators works without it, but any
call raises ``RuntimeError`` naming the missing file.  There is no CPU or
pure-torch fallback.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libteamtomo_amd.so")
_lock = threading.Lock()
_lib = None

ABI_VERSION = 13

TT_INTERP_NEAREST = 0
TT_INTERP_TRILINEAR = 1
TT_INTERP_BILINEAR = 1
TT_INTERP_BICUBIC = 2

TT_FLAG_DETERMINISTIC = 1
TT_FLAG_OVERWRITE = 2
TT_FLAG_PIXEL_ATOMIC = 4

TT_GATHER_AUTO = 0
TT_GATHER_PIXEL = 1
TT_GATHER_BRICK = 2

TT_BATCH_EXTRACT = 0
TT_BATCH_INSERT = 1

_p = ctypes.c_void_p
_i = ctypes.c_int
_l = ctypes.c_int64
_d = ctypes.c_double

# name -> argtypes (all return int status)
_SIGNATURES = {
    "tt_set_gather_path": [_i],
    "tt_set_scatter_path": [_i],
    "tt_plan_check": [_i, _i, _i, _i, _i, _i, _i, _l, _d, _p],
    "tt_extract_c64": [_p, _i, _i, _i, _i, _i, _i, _p, _l, _d, _d, _p, _p],
    "tt_extract_c64_noalloc": [_p, _i, _i, _i, _i, _i, _i, _p, _l, _d, _d, _p, _p],
    "tt_extract_mc_c64": [_p, _i, _i, _i, _i, _i, _i, _i, _p, _l, _d, _d, _p, _p],
    "tt_extract_bwd_c64": [_p, _p, _i, _i, _i, _i, _i, _i, _p, _l, _d, _d, _p, _p, _i, _p],
    "tt_insert_c64": [_p, _i, _i, _i, _p, _l, _d, _p, _p, _i, _p],
    "tt_insert_c128": [_p, _i, _i, _i, _p, _l, _d, _p, _p, _i, _p],
    "tt_insert_mc_c64": [_p, _i, _i, _i, _i, _p, _l, _d, _p, _p, _i, _p],
    "tt_insert_bwd_c64": [_p, _p, _p, _i, _i, _i, _p, _l, _d, _p, _p, _p],
    "tt_insert_bwd_c128": [_p, _p, _p, _i, _i, _i, _p, _l, _d, _p, _p, _p],
    "tt_sample3d_f32": [_p, _i, _i, _i, _i, _i, _p, _l, _i, _p, _p],
    "tt_sample3d_f64": [_p, _i, _i, _i, _i, _i, _p, _l, _i, _p, _p],
    "tt_sample3d_bwd_f32": [_p, _i, _i, _i, _i, _i, _p, _l, _i, _p, _p, _p, _p],
    "tt_sample3d_bwd_f64": [_p, _i, _i, _i, _i, _i, _p, _l, _i, _p, _p, _p, _p],
    "tt_splat3d_f32": [_p, _p, _l, _i, _i, _i, _i, _i, _i, _p, _p, _i, _p],
    "tt_splat3d_f64": [_p, _p, _l, _i, _i, _i, _i, _i, _i, _p, _p, _i, _p],
    "tt_extract2d_c64": [_p, _i, _i, _i, _i, _p, _l, _d, _p, _p],
    "tt_extract2d_bwd_c64": [_p, _p, _i, _i, _i, _i, _p, _l, _d, _p, _p, _p],
    "tt_sample2d_f32": [_p, _i, _i, _i, _i, _p, _l, _i, _p, _p],
    "tt_sample2d_f64": [_p, _i, _i, _i, _i, _p, _l, _i, _p, _p],
    "tt_splat2d_f32": [_p, _p, _l, _i, _i, _i, _i, _i, _p, _p, _i, _p],
    "tt_splat2d_f64": [_p, _p, _l, _i, _i, _i, _i, _i, _p, _p, _i, _p],
    "tt_affine3d_f32": [_p, _i, _i, _i, _i, _i, _p, _l, _i, _i, _i, _i, _i, _i, _i, _p, _p],
    "tt_affine2d_f32": [_p, _i, _i, _i, _i, _p, _l, _i, _i, _i, _i, _i, _p, _p],
    "tt_crop3d_f32": [_p, _i, _i, _i, _p, _l, _i, _p, _p, _p],
    "tt_crop2d_f32": [_p, _i, _i, _i, _p, _l, _i, _p, _p, _p],
    "tt_prepare_rotations_f32": [_p, _l, _i, _d, _p, _p],
    "tt_prepare_rotations_f64": [_p, _l, _i, _d, _p, _p],
    "tt_fourier_shift_c64": [_p, _l, _p, _l, _i, _i, _i, _i, _i, _i, _p, _p],
    "tt_fourier_shift_c128": [_p, _l, _p, _l, _i, _i, _i, _i, _i, _i, _p, _p],
    "tt_fourier_shift_bwd_c64": [_p, _p, _l, _p, _l, _i, _i, _i, _i, _i, _i, _p, _p, _p],
    "tt_fourier_shift_bwd_c128": [_p, _p, _l, _p, _l, _i, _i, _i, _i, _i, _i, _p, _p, _p],
    "tt_pad_sinc2_fftshift_f32": [_p, _i, _i, _p, _p],
    "tt_pad_sinc2_fftshift_f64": [_p, _i, _i, _p, _p],
    "tt_ifftshift_sinc2_crop_f32": [_p, _i, _i, _i, _p, _p],
    "tt_ifftshift_sinc2_crop_f64": [_p, _i, _i, _i, _p, _p],
    "tt_wdiv_ifftshift_c64": [_p, _p, _i, _i, _i, _p, _p],
    "tt_wdiv_ifftshift_c128": [_p, _p, _i, _i, _i, _p, _p],
    "tt_batch_create": [_i, _p, _i, _i, _p, ctypes.POINTER(ctypes.c_void_p)],
    "tt_batch_run": [_p, _p],
    "tt_batch_tile": [_p, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)],
    "tt_batch_destroy": [_p],
}


class BatchItem(ctypes.Structure):
    """``tt_batch_item`` of include/teamtomo_amd.h."""

    _fields_ = [
        ("input", _p),
        ("vol_d", _i), ("vol_h", _i), ("vol_wr", _i),
        ("img_d", _i), ("img_h", _i), ("img_w", _i),
        ("rot_zyx", _p),
        ("n_rot", _l),
        ("fftfreq_max", _d),
        ("x_tolerance", _d),
        ("output", _p),
        ("weights", _p),
    ]


def library_path() -> str:
    return _LIB_PATH


def load():
    """Load the shared library once and bind every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"teamtomo_amd: {_LIB_PATH} is missing; build it with "
                "`python -m teamtomo_amd._build` (hipcc, gfx950)"
            )
        lib = ctypes.CDLL(_LIB_PATH)
        lib.tt_last_error.restype = ctypes.c_char_p
        lib.tt_last_error.argtypes = []
        lib.tt_abi_version.restype = ctypes.c_int
        lib.tt_abi_version.argtypes = []
        if lib.tt_abi_version() != ABI_VERSION:
            raise RuntimeError(
                f"teamtomo_amd: {_LIB_PATH} has ABI {lib.tt_abi_version()}, expected {ABI_VERSION}; rebuild it"
            )
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = argtypes
        _lib = lib
    return _lib


def call(name: str, *args) -> None:
    """Call an entry point; raise RuntimeError with the library's message on failure."""
    lib = load()
    status = getattr(lib, name)(*args)
    if status != 0:
        msg = lib.tt_last_error().decode(errors="replace")
        kind = "invalid argument" if status == 1 else "HIP error"
        raise RuntimeError(f"teamtomo_amd.{name}: {kind}: {msg}")


def ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_of(t: torch.Tensor):
    """The current HIP stream of the tensor's device, as a void pointer."""
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def require_gpu(*tensors: torch.Tensor) -> None:
    for t in tensors:
        if t is not None and t.device.type != "cuda":
            raise RuntimeError(
                "teamtomo_amd runs on ROCm GPU tensors only (device 'cuda'); "
                f"got a tensor on '{t.device}'"
            )
