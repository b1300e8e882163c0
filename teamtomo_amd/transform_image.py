"""Drop-in replacements for torch_transform_image's affine warps.

Reference: packages/primitives/torch-transform-image/src/torch_transform_image/
transforms_3d.py:10-40 and transforms_2d.py:11-178.  The reference builds the
dense coordinate grid, multiplies it by the matrices and samples it with
sample_image_3d / sample_image_2d; here one HIP kernel computes each output
voxel's coordinate M (z, y, x, 1)^T and samples it (``tt_affine3d_f32``,
``tt_affine2d_f32``), so no grid is materialised.  The 3-D warp is
differentiable with respect to the image and the matrices: when either needs
a gradient it takes the reference's composition (coordinate grid, matmul,
then the differentiable sample_image_3d).  The 2-D warps are forward only: a
call whose image requires grad raises.
"""

from __future__ import annotations

import math
from typing import Literal, Optional

import torch

from . import _lib
from .image_interpolation import _mode, _no_grad, sample_image_3d

__all__ = [
    "affine_transform_image_3d",
    "affine_transform_image_2d",
    "rotate_then_shift_image_2d",
    "shift_then_rotate_image_2d",
]


def _check_image(image, ndim, name, grad_ok=False):
    _lib.require_gpu(image)
    if not grad_ok:
        _no_grad(image)
    if image.dtype not in (torch.float32, torch.complex64):
        # the reference samples with a float32 coordinate grid, which grid_sample
        # refuses for float64 / complex128 images
        raise RuntimeError(
            "grid_sampler(): expected input and grid to have same dtype, but input has "
            f"{torch.float64 if image.dtype in (torch.float64, torch.complex128) else image.dtype} "
            "and grid has torch.float32"
        )
    if image.ndim not in (ndim, ndim + 1):
        raise ValueError(f"{name}: image should have {ndim} or {ndim + 1} dimensions, got {tuple(image.shape)}")


def _matrices(matrices, device, grid_shape, flip):
    """float32 (..., k+1, k+1) matrices -> the (n, k, k+1) zyx / yx blocks laid
    out as (lead, m_1, ..., m_k), the leading batch shape and (m_1, ..., m_k):
    the matrices' batch dimensions broadcast against the output grid exactly
    as in the reference's ``matrices @ grid`` (transforms_3d.py:32-35)."""
    k = len(grid_shape)
    M = torch.as_tensor(matrices, device=device)
    if M.dtype != torch.float32:
        # torch.matmul of the matrices with the float32 grid needs float32 matrices
        raise RuntimeError(f"expected m1 and m2 to have the same dtype, but got: {M.dtype} != float")
    if M.shape[-2:] != (k + 1, k + 1):
        raise ValueError(f"matrices must have shape (..., {k + 1}, {k + 1}), got {tuple(M.shape)}")
    bshape = tuple(M.shape[:-2])
    full = torch.broadcast_shapes(bshape, tuple(grid_shape))  # the matmul's own broadcasting error
    lead = tuple(full[:-k])
    b = (1,) * (len(full) - len(bshape)) + bshape
    mgrid = tuple(b[-k:])
    M = M.reshape(b + (k + 1, k + 1)).expand(lead + mgrid + (k + 1, k + 1))
    if flip:  # xyz -> zyx (transforms_3d.py:24-29, transforms_2d.py:24-28)
        A = torch.flip(M[..., :k, :k], dims=(-2, -1))
        t = torch.flip(M[..., :k, k], dims=(-1,))
    else:
        A, t = M[..., :k, :k], M[..., :k, k]
    blk = torch.cat([A, t[..., None]], dim=-1).reshape(-1, k, k + 1).contiguous()
    n_lead = 1
    for v in lead:
        n_lead *= v
    return blk, lead, n_lead, mgrid


def affine_transform_image_3d(
    image: torch.Tensor,
    matrices: torch.Tensor,
    interpolation: Literal["nearest", "trilinear"],
    output_shape: Optional[tuple] = None,
    zyx_matrices: bool = False,
) -> torch.Tensor:
    """Warp a (d, h, w) or (c, d, h, w) image by (..., 4, 4) affine matrices.

    Reference: transforms_3d.py:10-40.  Output voxel (z, y, x) takes the
    sample at M (z, y, x, 1)^T (trilinear or nearest, 0 outside).  The
    matrices' batch dimensions broadcast against (d, h, w) as in the
    reference's matmul: a (4, 4) matrix gives (d, h, w), a (b, 1, 1, 1, 4, 4)
    stack gives (b, d, h, w); multichannel images add a trailing c.
    """
    _check_image(image, 3, "affine_transform_image_3d", grad_ok=True)
    d, h, w = (int(v) for v in (output_shape if output_shape else image.shape[-3:]))
    mode = _mode(interpolation)
    if torch.is_grad_enabled() and (image.requires_grad or (torch.is_tensor(matrices) and matrices.requires_grad)):
        return _affine_3d_differentiable(image, matrices, interpolation, (d, h, w), zyx_matrices)
    blk, lead, n, (md, mh, mw) = _matrices(matrices, image.device, (d, h, w), not zyx_matrices)
    multichannel = image.ndim == 4
    img = (image if multichannel else image.unsqueeze(0)).contiguous()
    C, D, H, W = img.shape
    out = torch.empty((n, d, h, w, C), dtype=image.dtype, device=image.device)
    _lib.call("tt_affine3d_f32", _lib.ptr(img), C, D, H, W, int(image.is_complex()), _lib.ptr(blk), n, md, mh, mw,
              d, h, w, mode, _lib.ptr(out), _lib.stream_of(image))
    out = out.reshape(*lead, d, h, w, C)
    return out if multichannel else out[..., 0]


def _affine_3d_differentiable(image, matrices, interpolation, shape, zyx_matrices):
    """The reference's composition (transforms_3d.py:22-40) in torch operations
    around the differentiable sampler: gradients reach the image through the
    sampler's backward kernel and the matrices through the matmul."""
    M = torch.as_tensor(matrices, device=image.device)
    if M.dtype != torch.float32:
        raise RuntimeError(f"expected m1 and m2 to have the same dtype, but got: {M.dtype} != float")
    if M.shape[-2:] != (4, 4):
        raise ValueError(f"matrices must have shape (..., 4, 4), got {tuple(M.shape)}")
    if not zyx_matrices:  # xyz -> zyx, out of place
        A = torch.flip(M[..., :3, :3], dims=(-2, -1))
        t = torch.flip(M[..., :3, 3:], dims=(-2,))
        M = torch.cat([torch.cat([A, t], dim=-1), M[..., 3:, :]], dim=-2)
    axes = [torch.arange(n, dtype=torch.float32, device=image.device) for n in shape]
    grid = torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=-1)  # (d, h, w, zyx)
    grid = torch.cat([grid, torch.ones(shape + (1,), device=image.device)], dim=-1)[..., None]
    grid = (M @ grid)[..., :3, 0]
    return sample_image_3d(image, grid, interpolation)


def affine_transform_image_2d(
    image: torch.Tensor,
    matrices: torch.Tensor,
    interpolation: Literal["nearest", "bilinear", "bicubic"],
    output_shape: Optional[tuple] = None,
    yx_matrices: bool = False,
) -> torch.Tensor:
    """Warp a (h, w) or (c, h, w) image by (..., 3, 3) affine matrices.

    Reference: transforms_2d.py:11-41.  Returns (..., h, w) or (..., h, w, c).
    """
    _check_image(image, 2, "affine_transform_image_2d")
    h, w = (int(v) for v in (output_shape if output_shape else image.shape[-2:]))
    mode = _mode(interpolation, linear="bilinear", bicubic=True)
    blk, lead, n, (mh, mw) = _matrices(matrices, image.device, (h, w), not yx_matrices)
    multichannel = image.ndim == 3
    img = (image if multichannel else image.unsqueeze(0)).contiguous()
    C, H, W = img.shape
    out = torch.empty((n, h, w, C), dtype=image.dtype, device=image.device)
    _lib.call("tt_affine2d_f32", _lib.ptr(img), C, H, W, int(image.is_complex()), _lib.ptr(blk), n, mh, mw, h, w,
              mode, _lib.ptr(out), _lib.stream_of(image))
    out = out.reshape(*lead, h, w, C)
    return out if multichannel else out[..., 0]


def _R2(angle, device):
    """xy rotation of homogeneous 2-D coordinates (torch-affine-utils transforms_2d.py R, yx=False)."""
    a = torch.deg2rad(torch.as_tensor(angle, dtype=torch.float32, device=device))
    c, s = torch.cos(a), torch.sin(a)
    M = torch.eye(3, device=device)
    M[0, 0], M[0, 1], M[1, 0], M[1, 1] = c, -s, s, c
    return M


def _T2(shift, device):
    """Translation of homogeneous 2-D coordinates (torch-affine-utils transforms_2d.py T)."""
    t = torch.atleast_1d(torch.as_tensor(shift, dtype=torch.float32, device=device))
    M = torch.eye(3, device=device)
    M[:2, 2] = t
    return M


def _rotate_shift_matrix(image, rotate, shift, rotate_first):
    device = image.device
    center = 0
    if rotate:
        h, w = image.shape[-2:]
        center = torch.tensor([h // 2, w // 2], dtype=torch.float32, device=device)
    center_t = torch.as_tensor(center, dtype=torch.float32, device=device)
    rotate_t = torch.as_tensor(rotate, dtype=torch.float32, device=device)
    # the shift is applied to the coordinate grid, so it is negated
    shift_t = -torch.as_tensor(shift, dtype=torch.float32, device=device)
    if rotate_t.numel() > 1 or shift_t.numel() > 2:
        raise NotImplementedError("Only single angle and single shift values are supported.")
    if rotate_first:
        return _T2(center_t, device) @ _R2(rotate_t, device) @ _T2(shift_t, device) @ _T2(-center_t, device)
    return _T2(center_t, device) @ _T2(shift_t, device) @ _R2(rotate_t, device) @ _T2(-center_t, device)


def rotate_then_shift_image_2d(
    image: torch.Tensor,
    rotate: int | float = 0,
    shift: list[float | int] | tuple[float | int, float | int] = (0, 0),
    interpolation: Literal["nearest", "bilinear", "bicubic"] = "bicubic",
) -> torch.Tensor:
    """Rotate about the image centre, then shift (transforms_2d.py:44-110)."""
    M = _rotate_shift_matrix(image, rotate, shift, rotate_first=True)
    return affine_transform_image_2d(image=image, matrices=M, interpolation=interpolation, yx_matrices=True)


def shift_then_rotate_image_2d(
    image: torch.Tensor,
    rotate: int | float = 0,
    shift: list[float | int] | tuple[float | int, float | int] = (0, 0),
    interpolation: Literal["nearest", "bilinear", "bicubic"] = "bicubic",
) -> torch.Tensor:
    """Shift, then rotate about the image centre (transforms_2d.py:113-178)."""
    M = _rotate_shift_matrix(image, rotate, shift, rotate_first=False)
    return affine_transform_image_2d(image=image, matrices=M, interpolation=interpolation, yx_matrices=True)
