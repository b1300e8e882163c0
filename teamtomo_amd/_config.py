"""Process-wide switches of the teamtomo_amd operators."""

from __future__ import annotations

import torch

_deterministic = False
_pixel_atomic = False


def set_deterministic(enabled: bool) -> None:
    """Require bitwise reproducible scatters (insertion and adjoints).

    The brick scatter accumulates in fixed point and is reproducible either
    way; with this set, the geometries only the pixel-atomic kernels handle
    (complex128 insertion, sizes beyond the brick engine) raise instead of
    running with float atomics."""
    global _deterministic
    _deterministic = bool(enabled)


def deterministic_enabled() -> bool:
    """True if set_deterministic(True) was called or torch's deterministic mode is on."""
    return _deterministic or torch.are_deterministic_algorithms_enabled()


def set_pixel_atomic(enabled: bool) -> None:
    """Use the pixel-driven global-atomic scatter instead of the brick-stationary
    LDS scatter (a cross-check path; slower at production sizes)."""
    global _pixel_atomic
    _pixel_atomic = bool(enabled)


def pixel_atomic_enabled() -> bool:
    return _pixel_atomic


_GATHER_PATHS = {"auto": 0, "pixel": 1, "brick": 2}


def set_gather_path(path: str) -> None:
    """Choose the extraction gather kernel: "auto" (default: the brick gather,
    or the pixel gather for geometries beyond the brick engine's limits),
    "pixel" (one lane per pixel, corners read through the caches) or "brick"
    (volume bricks staged in LDS, serving every slice crossing them).  The two
    agree to within 1e-6 of max|V| (the brick gather's FP32 frame coordinates
    against the pixel gather's FP64 ones), not bit for bit."""
    if path not in _GATHER_PATHS:
        raise ValueError(f"gather path must be one of {sorted(_GATHER_PATHS)}, got {path!r}")
    from . import _lib

    _lib.call("tt_set_gather_path", _GATHER_PATHS[path])


_SCATTER_PATHS = {"auto": 0, "brick": 1, "binned": 2}


def set_scatter_path(path: str) -> None:
    """Choose the scatter kernel of the complex64 insertion and extraction
    adjoint: "auto" (default: the binned scatter for calls with few slice
    pixels per volume brick, such as a tilt series into a large volume, the
    brick engine otherwise), "brick" (bricks find and walk the slice rows that
    cross them) or "binned" (slice pixels sorted into per-brick bins first).
    Both accumulate in fixed point and are bitwise reproducible; they agree to
    about 1e-6 relative, not bit for bit."""
    if path not in _SCATTER_PATHS:
        raise ValueError(f"scatter path must be one of {sorted(_SCATTER_PATHS)}, got {path!r}")
    from . import _lib

    _lib.call("tt_set_scatter_path", _SCATTER_PATHS[path])
