"""Process-wide switches of the teamtomo_amd operators."""

from __future__ import annotations

import torch

_deterministic = False


def set_deterministic(enabled: bool) -> None:
    """Select the fixed-order (bitwise reproducible) scatter for insertion and adjoints."""
    global _deterministic
    _deterministic = bool(enabled)


def deterministic_enabled() -> bool:
    """True if set_deterministic(True) was called or torch's deterministic mode is on."""
    return _deterministic or torch.are_deterministic_algorithms_enabled()
