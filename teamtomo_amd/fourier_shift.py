"""Drop-in replacements for torch_fourier_shift.

Reference: packages/primitives/torch-fourier-shift/src/torch_fourier_shift/
fourier_shift_dft.py:10-115 (a DFT times the phase grid of
phase_shift_grids.py:9-127) and fourier_shift_image.py:10-91 (rfftn, that
product, irfftn).  The phase grid is never materialised: one HIP kernel
multiplies each DFT element by exp(-2 pi i f . s) as it streams through
(tt_fourier_shift_* in include/teamtomo_amd.h), and one more pass gives the
backward (the DFT gradient and the shift gradient).  The FFTs of the image
variants stay on torch.
"""

from __future__ import annotations

import math

import torch

from . import _lib

__all__ = [
    "fourier_shift_dft_1d",
    "fourier_shift_dft_2d",
    "fourier_shift_dft_3d",
    "fourier_shift_image_1d",
    "fourier_shift_image_2d",
    "fourier_shift_image_3d",
]


def _result_dtype(dft: torch.Tensor, shifts: torch.Tensor) -> torch.dtype:
    """The reference's grid is float32 fftfreq times the shifts (float32 angles
    for integer or float32 shifts, float64 for float64 shifts) turned complex;
    the product with the DFT promotes."""
    angle = torch.promote_types(torch.float32, shifts.dtype)
    phase = torch.complex128 if angle == torch.float64 else torch.complex64
    out = torch.promote_types(dft.dtype, phase)
    return torch.complex128 if out == torch.complex128 else torch.complex64


def _launch(x, s, geom, out):
    ndim, n, per, stride, image_shape, rfft, fftshifted = geom
    d, h, w = (1,) * (3 - ndim) + tuple(image_shape)
    fn = "tt_fourier_shift_c64" if x.dtype == torch.complex64 else "tt_fourier_shift_c128"
    _lib.call(fn, _lib.ptr(x), stride, _lib.ptr(s), n, ndim, d, h, w, int(rfft), int(fftshifted), _lib.ptr(out),
              _lib.stream_of(x))
    return out


class _ShiftFn(torch.autograd.Function):
    """out[t] = dft[t or 0] * exp(-2 pi i f . shifts[t]) over n DFTs."""

    @staticmethod
    def forward(ctx, dft, shifts, geom):
        n, per = geom[1], geom[2]
        out = _launch(dft, shifts, geom, torch.empty((n, per), dtype=dft.dtype, device=dft.device))
        ctx.geom = geom
        ctx.save_for_backward(dft, shifts)
        return out

    @staticmethod
    def backward(ctx, grad):
        dft, shifts = ctx.saved_tensors
        ndim, n, per, stride, image_shape, rfft, fftshifted = ctx.geom
        d, h, w = (1,) * (3 - ndim) + tuple(image_shape)
        need_dft, need_s = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        grad = grad.to(dft.dtype).contiguous()
        grad_dft = torch.empty((n, per), dtype=dft.dtype, device=dft.device) if need_dft else None
        grad_s = torch.empty_like(shifts) if need_s else None
        fn = "tt_fourier_shift_bwd_c64" if dft.dtype == torch.complex64 else "tt_fourier_shift_bwd_c128"
        if need_dft or need_s:
            _lib.call(fn, _lib.ptr(grad), _lib.ptr(dft), stride, _lib.ptr(shifts), n, ndim, d, h, w, int(rfft),
                      int(fftshifted), _lib.ptr(grad_dft), _lib.ptr(grad_s), _lib.stream_of(grad))
        if need_dft and stride == 0:  # one DFT served every batch entry
            grad_dft = grad_dft.sum(0).reshape(dft.shape)
        elif need_dft:
            grad_dft = grad_dft.reshape(dft.shape)
        return grad_dft, grad_s, None


def _fourier_shift(dft, image_shape, shifts, rfft, fftshifted, ndim, inplace=False):
    """The DFT(s) times the phase ramp of the shift(s), batch dimensions
    broadcast as the reference's ``dft * phase_shifts`` does.  ``inplace``
    lets a caller that owns a fresh ``dft`` have it overwritten."""
    _lib.require_gpu(dft, shifts)
    if dft.device != shifts.device:
        raise RuntimeError(f"Expected all tensors to be on the same device, but found {dft.device} and {shifts.device}")
    image_shape = tuple(int(s) for s in image_shape)
    if len(image_shape) != ndim:
        raise ValueError(f"image_shape must have {ndim} entries, got {image_shape}")
    stored = image_shape[:-1] + (image_shape[-1] // 2 + 1,) if rfft else image_shape
    if ndim == 1:
        shifts = shifts.unsqueeze(-1)  # (...) -> (..., 1), phase_shift_grids.py:34
    elif shifts.ndim < 1 or shifts.shape[-1] != ndim:
        raise ValueError(f"shifts must have shape (..., {ndim}), got {tuple(shifts.shape)}")
    if dft.ndim < ndim or tuple(dft.shape[dft.ndim - ndim:]) != stored:
        raise RuntimeError(
            f"The size of the DFT {tuple(dft.shape)} does not match the phase grid {stored} of image_shape "
            f"{image_shape} (rfft={bool(rfft)})"
        )
    cdt = _result_dtype(dft, shifts)
    rdt = torch.float32 if cdt == torch.complex64 else torch.float64
    dbatch = tuple(dft.shape[: dft.ndim - ndim])
    batch = tuple(torch.broadcast_shapes(dbatch, tuple(shifts.shape[:-1])))
    n = math.prod(batch)
    per = math.prod(stored)
    if n == 0 or per == 0:
        return torch.empty(batch + stored, dtype=cdt, device=dft.device)
    s = shifts.to(rdt).expand(batch + (ndim,)).reshape(n, ndim).contiguous()
    x = dft.to(cdt)
    if dbatch == batch:
        x, stride = x.reshape(n, per).contiguous(), per
    elif math.prod(dbatch) == 1:
        x, stride = x.reshape(per).contiguous(), 0
    else:
        x, stride = x.expand(batch + stored).reshape(n, per).contiguous(), per
    geom = (ndim, n, per, stride, image_shape, bool(rfft), bool(fftshifted))
    if torch.is_grad_enabled() and (x.requires_grad or s.requires_grad):
        y = _ShiftFn.apply(x, s, geom)
    elif inplace and stride and x.data_ptr() == dft.data_ptr():
        y = _launch(x, s, geom, x)
    else:
        y = _launch(x, s, geom, torch.empty((n, per), dtype=cdt, device=x.device))
    return y.reshape(batch + stored)


def fourier_shift_dft_1d(dft: torch.Tensor, image_shape: tuple[int], shifts: torch.Tensor, rfft: bool,
                         fftshifted: bool) -> torch.Tensor:
    """(..., w) DFTs times the phase ramp of (...) shifts along w.

    Reference: fourier_shift_dft.py:10-42 (phase_shift_grids.py:9-45)."""
    return _fourier_shift(dft, image_shape, shifts, rfft, fftshifted, 1)


def fourier_shift_dft_2d(dft: torch.Tensor, image_shape: tuple[int, int], shifts: torch.Tensor, rfft: bool,
                         fftshifted: bool) -> torch.Tensor:
    """(..., h, w) DFTs times the phase ramp of (..., 2) (y, x) shifts.

    Reference: fourier_shift_dft.py:45-78 (phase_shift_grids.py:48-86)."""
    return _fourier_shift(dft, image_shape, shifts, rfft, fftshifted, 2)


def fourier_shift_dft_3d(dft: torch.Tensor, image_shape: tuple[int, int, int], shifts: torch.Tensor,
                         rfft: bool = False, fftshifted: bool = False) -> torch.Tensor:
    """(..., d, h, w) DFTs times the phase ramp of (..., 3) (z, y, x) shifts.

    Reference: fourier_shift_dft.py:81-115 (phase_shift_grids.py:89-127)."""
    return _fourier_shift(dft, image_shape, shifts, rfft, fftshifted, 3)


def _shift_image(image, shifts, ndim):
    dims = tuple(range(-ndim, 0))
    shape = tuple(image.shape[-ndim:])
    if image.numel() == 0:  # hipFFT refuses empty batches; the result is empty
        _lib.require_gpu(image, shifts)
        batch = torch.broadcast_shapes(tuple(image.shape[:-ndim]), tuple(shifts.shape[:-1] if ndim > 1 else shifts.shape))
        out_shape = tuple(batch) + shape[:-1] + (2 * (shape[-1] // 2),)
        return torch.empty(out_shape, dtype=torch.float32 if image.dtype != torch.float64 else torch.float64,
                           device=image.device)
    dft = torch.fft.rfftn(image, dim=dims)
    dft = _fourier_shift(dft, shape, shifts, True, False, ndim, inplace=True)
    # irfftn without s, as the reference calls it: an odd last axis comes back
    # one sample shorter
    return torch.real(torch.fft.irfftn(dft, dim=dims))


def fourier_shift_image_1d(image: torch.Tensor, shifts: torch.Tensor) -> torch.Tensor:
    """(..., w) images translated by (...) shifts (fourier_shift_image.py:10-35)."""
    return _shift_image(image, shifts, 1)


def fourier_shift_image_2d(image: torch.Tensor, shifts: torch.Tensor) -> torch.Tensor:
    """(..., h, w) images translated by (..., 2) shifts (fourier_shift_image.py:38-63)."""
    return _shift_image(image, shifts, 2)


def fourier_shift_image_3d(image: torch.Tensor, shifts: torch.Tensor) -> torch.Tensor:
    """(..., d, h, w) images translated by (..., 3) shifts (fourier_shift_image.py:66-91)."""
    return _shift_image(image, shifts, 3)
