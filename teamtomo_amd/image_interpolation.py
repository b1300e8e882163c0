"""Drop-in replacements for torch_image_interpolation's 3-D, 2-D and 1-D entry points.

Reference: packages/primitives/torch-image-interpolation/src/
torch_image_interpolation/image_interpolation_3d.py, image_interpolation_2d.py
and image_interpolation_1d.py.
Sampling and splatting run as gfx950 HIP kernels through the C ABI
(include/teamtomo_amd.h).  ``sample_image_3d`` is differentiable with respect
to the image and the coordinates, as the reference's grid_sample is; its
backward is a HIP kernel (tt_sample3d_bwd_*).  The other samplers and the
splats are forward-only: a call whose inputs require grad raises, rather than
returning a result without a gradient.
"""

from __future__ import annotations

from typing import Literal

import torch

from . import _lib

__all__ = [
    "sample_image_3d",
    "insert_into_image_3d",
    "sample_image_2d",
    "insert_into_image_2d",
    "sample_image_1d",
    "insert_into_image_1d",
]

_REAL_OF = {
    torch.float32: torch.float32,
    torch.complex64: torch.float32,
    torch.float64: torch.float64,
    torch.complex128: torch.float64,
}


def _mode(interpolation: str, linear: str = "trilinear", bicubic: bool = False) -> int:
    if interpolation == linear:
        return _lib.TT_INTERP_TRILINEAR
    if interpolation == "nearest":
        return _lib.TT_INTERP_NEAREST
    if bicubic and interpolation == "bicubic":
        return _lib.TT_INTERP_BICUBIC
    # grid_sample's own error for an unknown mode
    raise ValueError(
        "nn.functional.grid_sample(): expected mode to be 'bilinear', 'nearest' or "
        f"'bicubic', but got: '{interpolation}'"
    )


def _no_grad(*tensors) -> None:
    if torch.is_grad_enabled() and any(t.requires_grad for t in tensors):
        raise RuntimeError(
            "teamtomo_amd: the generic samplers have no backward; call them on tensors "
            "that do not require grad (or under torch.no_grad())"
        )


class _Sample3dFn(torch.autograd.Function):
    """samples (n, c) of a (c, d, h, w) image at (n, 3) coordinates; backward:
    the image gradient (transposed trilinear weights) and the coordinate
    gradient (analytic weight derivatives), one kernel."""

    @staticmethod
    def forward(ctx, img, crd, mode):
        C, d, h, w = img.shape
        n = crd.shape[0]
        real_t = _REAL_OF[img.dtype]
        out = torch.empty((n, C), dtype=img.dtype, device=img.device)
        fn = "tt_sample3d_f32" if real_t == torch.float32 else "tt_sample3d_f64"
        _lib.call(fn, _lib.ptr(img), C, d, h, w, int(img.is_complex()), _lib.ptr(crd), n, mode, _lib.ptr(out),
                  _lib.stream_of(img))
        ctx.save_for_backward(img, crd)
        ctx.mode = mode
        return out

    @staticmethod
    def backward(ctx, grad_out):
        img, crd = ctx.saved_tensors
        need_i, need_c = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_i or need_c):
            return None, None, None
        C, d, h, w = img.shape
        n = crd.shape[0]
        real_t = _REAL_OF[img.dtype]
        g = grad_out.to(img.dtype).contiguous()
        gi = torch.zeros_like(img) if need_i else None
        gc = torch.empty_like(crd) if need_c else None
        fn = "tt_sample3d_bwd_f32" if real_t == torch.float32 else "tt_sample3d_bwd_f64"
        _lib.call(fn, _lib.ptr(img), C, d, h, w, int(img.is_complex()), _lib.ptr(crd), n, ctx.mode, _lib.ptr(g),
                  _lib.ptr(gi), _lib.ptr(gc), _lib.stream_of(img))
        return gi, gc, None


def sample_image_3d(
    image: torch.Tensor,
    coordinates: torch.Tensor,
    interpolation: Literal["nearest", "trilinear"] = "trilinear",
) -> torch.Tensor:
    """Sample a (d, h, w) or (c, d, h, w) image at (..., 3) zyx array coordinates.

    Reference: image_interpolation_3d.py:11-101.  Returns (...) or (..., c);
    points outside [0, N-1] in any dimension give 0.  Differentiable with
    respect to ``image`` and ``coordinates`` (grid_sample's gradients with
    border padding: 0 along an axis where the coordinate sits on or beyond
    the border, 0 for samples outside the image, no coordinate gradient for
    nearest sampling).
    """
    if image.ndim not in (3, 4):
        raise ValueError(f"image should have shape (d, h, w) or (c, d, h, w), got {image.shape}")
    _lib.require_gpu(image)
    if image.dtype not in _REAL_OF:
        raise TypeError(f"sample_image_3d: unsupported image dtype {image.dtype}")
    real_t = _REAL_OF[image.dtype]
    if coordinates.is_floating_point() and coordinates.dtype != real_t:
        raise RuntimeError(
            "grid_sampler(): expected input and grid to have same dtype, but input has "
            f"{real_t} and grid has {coordinates.dtype}"
        )
    mode = _mode(interpolation)
    multichannel = image.ndim == 4
    img = image if multichannel else image.unsqueeze(0)
    lead = tuple(coordinates.shape[:-1])
    crd = coordinates.to(device=image.device, dtype=real_t).reshape(-1, 3).contiguous()
    out = _Sample3dFn.apply(img.contiguous(), crd, mode)
    out = out.reshape(*lead, img.shape[0])
    return out if multichannel else out[..., 0]


def insert_into_image_3d(
    values: torch.Tensor,
    coordinates: torch.Tensor,
    image: torch.Tensor,
    weights: torch.Tensor | None = None,
    interpolation: Literal["nearest", "trilinear"] = "trilinear",
) -> tuple[torch.Tensor, torch.Tensor]:
    """Splat values into ``image`` (and ``weights``) in place; returns both.

    Reference: image_interpolation_3d.py:104-275.  Trilinear splats use
    floor/ceil corners, nearest rounds half to even with weight 1; points
    outside [0, N-1] are dropped and an unknown mode changes nothing.
    """
    multichannel = image.ndim == 4
    d, h, w = image.shape[-3:]
    values_shape = values.shape[:-1] if multichannel else values.shape
    if values_shape != coordinates.shape[:-1]:
        raise ValueError("One coordinate triplet is required for each value in data.")
    if coordinates.shape[-1] != 3:
        raise ValueError("Coordinates must be 3D with shape (..., 3).")
    if image.dtype != values.dtype:
        raise ValueError("Image and values must have the same dtype.")
    _lib.require_gpu(image, values)
    _no_grad(image, values)
    if weights is None:
        weights = torch.zeros(size=(d, h, w), dtype=torch.float32, device=image.device)
    if interpolation not in ("nearest", "trilinear"):
        return image, weights
    if image.dtype not in _REAL_OF:
        raise TypeError(f"insert_into_image_3d: unsupported image dtype {image.dtype}")
    if weights.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"insert_into_image_3d: weights must be float32 or float64, got {weights.dtype}")
    if not image.is_contiguous() or not weights.is_contiguous():
        raise ValueError("insert_into_image_3d: image and weights must be contiguous (updated in place)")
    C = image.shape[0] if multichannel else 1
    vals = values.reshape(-1, C).contiguous()
    crd = coordinates.to(device=image.device, dtype=torch.float32).reshape(-1, 3).contiguous()
    n = crd.shape[0]
    real_t = _REAL_OF[image.dtype]
    fn = "tt_splat3d_f32" if real_t == torch.float32 else "tt_splat3d_f64"
    _lib.call(fn, _lib.ptr(vals), _lib.ptr(crd), n, C, int(image.is_complex()), d, h, w,
              _lib.TT_INTERP_TRILINEAR if interpolation == "trilinear" else _lib.TT_INTERP_NEAREST,
              _lib.ptr(image), _lib.ptr(weights), int(weights.dtype == torch.float64),
              _lib.stream_of(image))
    return image, weights


def sample_image_2d(
    image: torch.Tensor,
    coordinates: torch.Tensor,
    interpolation: Literal["nearest", "bilinear", "bicubic"] = "bilinear",
) -> torch.Tensor:
    """Sample a (h, w) or (c, h, w) image at (..., 2) yx array coordinates.

    Reference: image_interpolation_2d.py:11-101.  Returns (...) or (..., c);
    points outside [0, N-1] in any dimension give 0.  bicubic is grid_sample's
    Keys cubic convolution with border-clipped taps.
    """
    if image.ndim not in (2, 3):
        raise ValueError(f"image should have shape (h, w) or (c, h, w), got {image.shape}")
    _lib.require_gpu(image)
    _no_grad(image, coordinates)
    if image.dtype not in _REAL_OF:
        raise TypeError(f"sample_image_2d: unsupported image dtype {image.dtype}")
    real_t = _REAL_OF[image.dtype]
    if coordinates.is_floating_point() and coordinates.dtype != real_t:
        raise RuntimeError(
            "grid_sampler(): expected input and grid to have same dtype, but input has "
            f"{real_t} and grid has {coordinates.dtype}"
        )
    mode = _mode(interpolation, linear="bilinear", bicubic=True)
    multichannel = image.ndim == 3
    img = (image if multichannel else image.unsqueeze(0)).contiguous()
    C, h, w = img.shape
    lead = tuple(coordinates.shape[:-1])
    crd = coordinates.to(device=image.device, dtype=real_t).reshape(-1, 2).contiguous()
    n = crd.shape[0]
    out = torch.empty((n, C), dtype=image.dtype, device=image.device)
    fn = "tt_sample2d_f32" if real_t == torch.float32 else "tt_sample2d_f64"
    _lib.call(fn, _lib.ptr(img), C, h, w, int(image.is_complex()), _lib.ptr(crd), n, mode, _lib.ptr(out),
              _lib.stream_of(image))
    out = out.reshape(*lead, C)
    return out if multichannel else out[..., 0]


def insert_into_image_2d(
    values: torch.Tensor,
    coordinates: torch.Tensor,
    image: torch.Tensor,
    weights: torch.Tensor | None = None,
    interpolation: Literal["nearest", "bilinear"] = "bilinear",
) -> tuple[torch.Tensor, torch.Tensor]:
    """Splat values into ``image`` (and ``weights``) in place; returns both.

    Reference: image_interpolation_2d.py:104-275.  Bilinear splats use
    floor/ceil corners, nearest rounds half to even with weight 1; points
    outside [0, N-1] are dropped and an unknown mode changes nothing.
    """
    multichannel = image.ndim == 3
    h, w = image.shape[-2:]
    values_shape = values.shape[:-1] if multichannel else values.shape
    if values_shape != coordinates.shape[:-1]:
        raise ValueError("One coordinate pair is required for each value in data.")
    if coordinates.shape[-1] != 2:
        raise ValueError("Coordinates must be 2D with shape (..., 2).")
    if image.dtype != values.dtype:
        raise ValueError("Image and values must have the same dtype.")
    _lib.require_gpu(image, values)
    _no_grad(image, values)
    if weights is None:
        weights = torch.zeros(size=(h, w), dtype=torch.float32, device=image.device)
    if interpolation not in ("nearest", "bilinear"):
        return image, weights
    if image.dtype not in _REAL_OF:
        raise TypeError(f"insert_into_image_2d: unsupported image dtype {image.dtype}")
    if weights.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"insert_into_image_2d: weights must be float32 or float64, got {weights.dtype}")
    if not image.is_contiguous() or not weights.is_contiguous():
        raise ValueError("insert_into_image_2d: image and weights must be contiguous (updated in place)")
    C = image.shape[0] if multichannel else 1
    vals = values.reshape(-1, C).contiguous()
    crd = coordinates.to(device=image.device, dtype=torch.float32).reshape(-1, 2).contiguous()
    n = crd.shape[0]
    fn = "tt_splat2d_f32" if _REAL_OF[image.dtype] == torch.float32 else "tt_splat2d_f64"
    _lib.call(fn, _lib.ptr(vals), _lib.ptr(crd), n, C, int(image.is_complex()), h, w,
              _lib.TT_INTERP_BILINEAR if interpolation == "bilinear" else _lib.TT_INTERP_NEAREST,
              _lib.ptr(image), _lib.ptr(weights), int(weights.dtype == torch.float64),
              _lib.stream_of(image))
    return image, weights


def sample_image_1d(
    image: torch.Tensor,
    coordinates: torch.Tensor,
    interpolation: Literal["nearest", "linear", "cubic"] = "linear",
) -> torch.Tensor:
    """Sample a (w,) or (c, w) image at (...) array coordinates.

    Reference: image_interpolation_1d.py:11-107, which runs grid_sample on the
    image repeated to height 2 at y = 0 with border padding.  Here the same
    composition runs on the 2-D sampling kernel: every y tap reads the one row,
    so linear, cubic (Keys, border-clipped taps) and nearest are the 1-D rules.
    Returns (...) or (..., c); points outside [0, w-1] give 0.
    """
    if image.ndim not in (1, 2):
        raise ValueError(f"image should have shape (w, ) or (c, w), got {image.shape}")
    mode = {"linear": "bilinear", "cubic": "bicubic"}.get(interpolation, interpolation)
    multichannel = image.ndim == 2
    img = (image if multichannel else image.unsqueeze(0)).unsqueeze(1).expand(-1, 2, -1)
    crd = torch.stack([torch.zeros_like(coordinates), coordinates], dim=-1)
    out = sample_image_2d(img, crd, interpolation=mode)  # (..., c)
    return out if multichannel else out[..., 0]


def insert_into_image_1d(
    values: torch.Tensor,
    coordinates: torch.Tensor,
    image: torch.Tensor,
    weights: torch.Tensor | None = None,
    interpolation: Literal["nearest", "linear"] = "linear",
) -> tuple[torch.Tensor, torch.Tensor]:
    """Splat (...) or (..., c) values at (...) coordinates into a (w,) or
    (c, w) ``image`` and (w,) ``weights``, in place; returns both.

    Reference: image_interpolation_1d.py:110-244.  Linear splats use floor/ceil
    corners with weights 1 - t and t, nearest rounds half to even with weight
    1; points outside [0, w-1] are dropped and an unknown mode changes
    nothing.  Runs on the 2-D splat kernel with the image seen as (c, 1, w).
    """
    multichannel = image.ndim == 2
    w = image.shape[-1]
    values_shape = values.shape[:-1] if multichannel else values.shape
    if values_shape != coordinates.shape:
        raise ValueError("One coordinate is required for each value in data.")
    if image.dtype != values.dtype:
        raise ValueError("Image and values must have the same dtype.")
    _lib.require_gpu(image, values)
    if weights is None:
        weights = torch.zeros(size=(w,), dtype=torch.float32, device=image.device)
    if interpolation not in ("nearest", "linear"):
        return image, weights
    if not image.is_contiguous() or not weights.is_contiguous():
        raise ValueError("insert_into_image_1d: image and weights must be contiguous (updated in place)")
    C = image.shape[0] if multichannel else 1
    vals = values if multichannel else values.unsqueeze(-1)
    crd = torch.stack([torch.zeros_like(coordinates, dtype=torch.float32), coordinates.to(torch.float32)], dim=-1)
    insert_into_image_2d(vals, crd, image.view(C, 1, w), weights.view(1, w),
                         interpolation="bilinear" if interpolation == "linear" else "nearest")
    return image, weights
