"""Process-wide switches of the teamtomo_amd operators."""

from __future__ import annotations

import torch

_deterministic = False
_pixel_atomic = False


def set_deterministic(enabled: bool) -> None:
    """Select the fixed-order (bitwise reproducible) scatter for insertion and adjoints."""
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
