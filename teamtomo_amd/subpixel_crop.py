"""Drop-in replacements for torch_subpixel_crop.

Reference: packages/primitives/torch-subpixel-crop/src/torch_subpixel_crop/
subpixel_crop_3d.py:12-108 and subpixel_crop_2d.py:12-163.  Patches are cut
at round(position) (zeros outside the image) by one HIP kernel that also
applies a (p, p[, p]) mask, and moved by the subpixel remainder with the
Fourier-shift kernel on the rfft (torch-fourier-shift phase_shift_grids.py:60-127);
the FFTs stay on torch.  Forward only: an input that requires grad raises.
"""

from __future__ import annotations

from typing import Optional

import torch

from . import _lib
from .fourier_shift import _fourier_shift
from .image_interpolation import _no_grad

__all__ = ["subpixel_crop_3d", "subpixel_crop_2d"]


def _check(image, name):
    _lib.require_gpu(image)
    _no_grad(image)
    if image.dtype != torch.float32:
        # the reference samples with a float32 grid, which grid_sample refuses
        # for other image dtypes
        raise RuntimeError(
            f"grid_sampler(): expected input and grid to have same dtype, but input has {image.dtype} "
            "and grid has torch.float32"
        )


def _fused_mask(mask, shape, device):
    """The mask as a flat (p...) float32 buffer when it has exactly the patch
    shape (applied in the crop kernel), else None (multiplied afterwards)."""
    if mask is None:
        return None
    m = torch.as_tensor(mask, device=device)
    if tuple(m.shape) == tuple(shape) and m.dtype == torch.float32:
        return m.contiguous()
    return None


def _shift(patches, shifts, ndim, p, return_rfft, decenter):
    """The subpixel shift (fourier_shift_image_nd, or the rfft with
    fourier_shift_dft_nd when return_rfft), as the reference composes it."""
    dims = tuple(range(-ndim, 0))
    if patches.numel() == 0:  # hipFFT refuses empty batches; the result is empty
        lead = tuple(patches.shape[:-ndim]) + (p,) * (ndim - 1)
        if return_rfft:
            return torch.empty(lead + (p // 2 + 1,), dtype=torch.complex64, device=patches.device)
        return torch.empty(lead + (2 * (p // 2),), dtype=torch.float32, device=patches.device)
    dft = torch.fft.rfftn(patches, dim=dims)
    if return_rfft and decenter:
        shifts = shifts + p / 2
    dft = _fourier_shift(dft, (p,) * ndim, shifts.to(torch.float32), True, False, ndim, inplace=True)
    if return_rfft:
        return dft
    return torch.real(torch.fft.irfftn(dft, dim=dims))


def subpixel_crop_3d(
    image: torch.Tensor,
    positions: torch.Tensor,
    sidelength: int,
    mask: Optional[torch.Tensor] = None,
    return_rfft: bool = False,
    decenter: bool = False,
) -> torch.Tensor:
    """Cubic patches of a (d, h, w) volume centred at (..., 3) zyx positions
    with subpixel precision (subpixel_crop_3d.py:12-108)."""
    _check(image, "subpixel_crop_3d")
    d, h, w = image.shape
    pos = torch.as_tensor(positions, device=image.device)
    lead = tuple(pos.shape[:-1])
    pos = pos.reshape(-1, 3)
    ipos = torch.round(pos)
    shifts = ipos - pos
    p = int(sidelength)
    fm = _fused_mask(mask, (p, p, p), image.device)
    patches = torch.empty((pos.shape[0], p, p, p), dtype=torch.float32, device=image.device)
    _lib.call("tt_crop3d_f32", _lib.ptr(image.contiguous()), d, h, w, _lib.ptr(ipos.to(torch.int32).contiguous()),
              pos.shape[0], p, _lib.ptr(fm), _lib.ptr(patches), _lib.stream_of(image))
    if mask is not None and fm is None:
        patches = patches * mask
    out = _shift(patches, shifts, 3, p, return_rfft, decenter)
    return out.reshape(*lead, *out.shape[-3:])


def subpixel_crop_2d(
    image: torch.Tensor,
    positions: torch.Tensor,
    sidelength: int,
    mask: Optional[torch.Tensor] = None,
    return_rfft: bool = False,
    decenter: bool = False,
) -> torch.Tensor:
    """Square patches of (b, h, w) or (h, w) images centred at (..., b, 2) or
    (..., 2) yx positions with subpixel precision (subpixel_crop_2d.py:12-163)."""
    _check(image, "subpixel_crop_2d")
    batched = image.ndim == 3
    imgs = image if batched else image[None]
    pos = torch.as_tensor(positions, device=image.device)
    if not batched:
        pos = pos[..., None, :]
    B, h, w = imgs.shape
    lead = tuple(pos.shape[:-2])
    pos = pos.reshape(-1, pos.shape[-2], 2)
    if pos.shape[1] != B:
        raise ValueError("Mismatch in batch size for images and positions.")
    ipos = torch.round(pos)
    shifts = ipos - pos
    p = int(sidelength)
    fm = _fused_mask(mask, (p, p), image.device)
    patches = torch.empty((pos.shape[0], B, p, p), dtype=torch.float32, device=image.device)
    _lib.call("tt_crop2d_f32", _lib.ptr(imgs.contiguous()), B, h, w, _lib.ptr(ipos.to(torch.int32).contiguous()),
              pos.shape[0], p, _lib.ptr(fm), _lib.ptr(patches), _lib.stream_of(image))
    if mask is not None and fm is None:
        patches = patches * mask
    out = _shift(patches, shifts, 2, p, return_rfft, decenter)
    out = out.reshape(*lead, B, *out.shape[-2:])
    return out if batched else out[..., 0, :, :]
