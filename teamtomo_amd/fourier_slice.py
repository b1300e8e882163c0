"""Drop-in replacements for torch_fourier_slice's 3-D central-slice operators.

Same names, arguments, return values and errors as the reference package
(packages/primitives/torch-fourier-slice/src/torch_fourier_slice); the
gather and scatter run as gfx950 HIP kernels through the C ABI in
``include/teamtomo_amd.h``, FFTs and elementwise pre/post-processing stay on
torch.  Differentiation uses hand-written backward kernels: the volume
gradient of an extraction is the adjoint scatter, the rotation gradients come
from analytic trilinear-weight derivatives, and the slice gradient of an
insertion is the adjoint gather.  Rotated coordinates are recomputed in the
backward pass, never stored.

Deliberate differences from the reference (documented, not bugs):
* the caller's float32 ``rotation_matrices`` are not modified in place by the
  near-zero clamp (the reference writes zeros into them,
  slice_extraction/_extract_central_slices_rfft_3d.py:147-151), and a float32
  leaf with ``requires_grad`` works instead of raising;
* coordinates are evaluated in FP64 (the reference uses float32), so results
  sit closer to the exact values.
"""

from __future__ import annotations

import torch
import torch.nn.functional as F

from . import _config, _lib
from ._config import deterministic_enabled

__all__ = [
    "extract_central_slices_rfft_3d",
    "extract_central_slices_rfft_3d_multichannel",
    "insert_central_slices_rfft_3d",
    "insert_central_slices_rfft_3d_multichannel",
    "project_3d_to_2d",
    "project_3d_to_2d_multichannel",
    "backproject_2d_to_3d",
    "backproject_2d_to_3d_multichannel",
    "extract_central_slices_rfft_2d",
    "project_2d_to_1d",
]


def _rfft_width(w: int) -> int:
    return int(w / 2 + 1)


def _fmax_arg(fftfreq_max) -> float:
    return -1.0 if fftfreq_max is None else float(fftfreq_max)


def _scatter_flags() -> int:
    """Scatter outputs are freshly allocated, so they are written, not accumulated."""
    flags = _lib.TT_FLAG_OVERWRITE
    if deterministic_enabled():
        flags |= _lib.TT_FLAG_DETERMINISTIC
    if _config.pixel_atomic_enabled():
        flags |= _lib.TT_FLAG_PIXEL_ATOMIC
    return flags


def _prepare_rotations(rotation_matrices, device, zyx_matrices, rot_tolerance):
    """float32 cast, near-zero clamp (out of place) and xyz->zyx flip.

    slice_extraction/_extract_central_slices_rfft_3d.py:147-154;
    slice_insertion/_insert_central_slices_rfft_3d.py:16,66-67.
    """
    Rin = rotation_matrices
    if (Rin.device == device and device.type == "cuda" and Rin.dtype in (torch.float32, torch.float64)
            and Rin.ndim >= 2 and tuple(Rin.shape[-2:]) == (3, 3)
            and not (torch.is_grad_enabled() and Rin.requires_grad)):
        # one kernel for the cast, the clamp and the flip (no autograd needed)
        src = Rin.contiguous()
        out = torch.empty(Rin.shape, dtype=torch.float32, device=device)
        fn = "tt_prepare_rotations_f32" if src.dtype == torch.float32 else "tt_prepare_rotations_f64"
        _lib.call(fn, _lib.ptr(src), src.numel() // 9, int(not zyx_matrices),
                  -1.0 if rot_tolerance is None else float(rot_tolerance), _lib.ptr(out), _lib.stream_of(out))
        return out
    R = rotation_matrices.to(device=device, dtype=torch.float32)
    if rot_tolerance is not None:
        R = torch.where(torch.abs(R) < rot_tolerance, torch.zeros((), dtype=R.dtype, device=device), R)
    if not zyx_matrices:
        R = torch.flip(R, dims=(-2, -1))
    return R


class _ExtractFn(torch.autograd.Function):
    """slices = gather(volume, R); backward = adjoint scatter + rotation grad."""

    @staticmethod
    def forward(ctx, volume, R, image_shape, fftfreq_max, x_tolerance):
        n = R.shape[0]
        D, H, Wr = volume.shape
        img_d, img_h, img_w = (int(s) for s in image_shape)
        out = torch.empty((n, img_h, _rfft_width(img_w)), dtype=volume.dtype, device=volume.device)
        _lib.call(
            "tt_extract_c64", _lib.ptr(volume), D, H, Wr, img_d, img_h, img_w, _lib.ptr(R), n,
            _fmax_arg(fftfreq_max), float(x_tolerance), _lib.ptr(out), _lib.stream_of(volume),
        )
        ctx.save_for_backward(volume, R)
        ctx.geom = (img_d, img_h, img_w, fftfreq_max, x_tolerance)
        return out

    @staticmethod
    def backward(ctx, grad_slices):
        volume, R = ctx.saved_tensors
        img_d, img_h, img_w, fftfreq_max, x_tolerance = ctx.geom
        need_v, need_r = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_v or need_r):
            return None, None, None, None, None
        g = grad_slices.to(torch.complex64).contiguous()
        D, H, Wr = volume.shape
        gV = torch.empty_like(volume) if need_v else None
        gR = torch.zeros((R.shape[0], 3, 3), dtype=torch.float64, device=R.device) if need_r else None
        _lib.call(
            "tt_extract_bwd_c64", _lib.ptr(g), _lib.ptr(volume), D, H, Wr, img_d, img_h, img_w,
            _lib.ptr(R), R.shape[0], _fmax_arg(fftfreq_max), float(x_tolerance), _lib.ptr(gV),
            _lib.ptr(gR), _scatter_flags(), _lib.stream_of(volume),
        )
        return gV, (gR.to(R.dtype) if gR is not None else None), None, None, None


def extract_central_slices_rfft_3d(
    volume_rfft: torch.Tensor,
    image_shape: tuple[int, int, int],
    rotation_matrices: torch.Tensor,
    fftfreq_max: float | None = None,
    zyx_matrices: bool = False,
    x_tolerance: float = 1e-8,
    rot_tolerance: float | None = 1e-8,
) -> torch.Tensor:
    """Extract central slices from an fftshifted rfft volume.

    Reference: slice_extraction/_extract_central_slices_rfft_3d.py:84-222.
    ``volume_rfft`` is (d, h, w//2+1) complex64 on a ROCm device; returns
    (..., h, w//2+1) for rotation matrices of shape (..., 3, 3).
    """
    assert volume_rfft.ndim == 3, "Input volume_rfft must be 3D tensor."
    _lib.require_gpu(volume_rfft)
    if volume_rfft.dtype != torch.complex64:
        raise RuntimeError(
            "extract_central_slices_rfft_3d: volume_rfft must be complex64 "
            f"(grid_sample needs input and grid of the same dtype), got {volume_rfft.dtype}"
        )
    R = _prepare_rotations(rotation_matrices, volume_rfft.device, zyx_matrices, rot_tolerance)
    stack_shape = tuple(R.shape[:-2])
    R = R.reshape(-1, 3, 3).contiguous()
    out = _ExtractFn.apply(volume_rfft.contiguous(), R, tuple(image_shape), fftfreq_max, x_tolerance)
    return out.reshape(*stack_shape, *out.shape[-2:])


class _ExtractMCFn(torch.autograd.Function):
    """slices (n, c, h, w) = gather of a (c, d, h, w) volume, two channels per
    brick launch; backward = per-channel adjoint scatter + rotation grad."""

    @staticmethod
    def forward(ctx, volume, R, image_shape, fftfreq_max, x_tolerance):
        n = R.shape[0]
        C, D, H, Wr = volume.shape
        img_d, img_h, img_w = (int(s) for s in image_shape)
        out = torch.empty((n, C, img_h, _rfft_width(img_w)), dtype=volume.dtype, device=volume.device)
        _lib.call(
            "tt_extract_mc_c64", _lib.ptr(volume), C, D, H, Wr, img_d, img_h, img_w, _lib.ptr(R), n,
            _fmax_arg(fftfreq_max), float(x_tolerance), _lib.ptr(out), _lib.stream_of(volume),
        )
        ctx.save_for_backward(volume, R)
        ctx.geom = (img_d, img_h, img_w, fftfreq_max, x_tolerance)
        return out

    @staticmethod
    def backward(ctx, grad_slices):
        volume, R = ctx.saved_tensors
        img_d, img_h, img_w, fftfreq_max, x_tolerance = ctx.geom
        need_v, need_r = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_v or need_r):
            return None, None, None, None, None
        C, D, H, Wr = volume.shape
        gV = torch.empty_like(volume) if need_v else None
        gR = torch.zeros((R.shape[0], 3, 3), dtype=torch.float64, device=R.device) if need_r else None
        for c in range(C):
            g = grad_slices[:, c].to(torch.complex64).contiguous()
            _lib.call(
                "tt_extract_bwd_c64", _lib.ptr(g), _lib.ptr(volume[c]), D, H, Wr, img_d, img_h, img_w,
                _lib.ptr(R), R.shape[0], _fmax_arg(fftfreq_max), float(x_tolerance),
                _lib.ptr(gV[c]) if need_v else None, _lib.ptr(gR), _scatter_flags(), _lib.stream_of(volume),
            )
        return gV, (gR.to(R.dtype) if gR is not None else None), None, None, None


def extract_central_slices_rfft_3d_multichannel(
    volume_rfft: torch.Tensor,
    image_shape: tuple[int, int, int],
    rotation_matrices: torch.Tensor,
    fftfreq_max: float | None = None,
    zyx_matrices: bool = False,
    x_tolerance: float = 1e-8,
    rot_tolerance: float | None = 1e-8,
) -> torch.Tensor:
    """Multichannel extraction: (c, d, h, w//2+1) -> (..., c, h, w//2+1).

    Reference: slice_extraction/_extract_central_slices_rfft_3d.py:225-367.
    All channels share the geometry; the brick gather serves two channels per
    launch, so the slice pixels are enumerated once per pair of channels.
    """
    assert volume_rfft.ndim == 4, "Input volume_rfft must be 4D tensor."
    _lib.require_gpu(volume_rfft)
    if volume_rfft.dtype != torch.complex64:
        raise RuntimeError(
            "extract_central_slices_rfft_3d_multichannel: volume_rfft must be complex64 "
            f"(grid_sample needs input and grid of the same dtype), got {volume_rfft.dtype}"
        )
    R = _prepare_rotations(rotation_matrices, volume_rfft.device, zyx_matrices, rot_tolerance)
    stack_shape = tuple(R.shape[:-2])
    R = R.reshape(-1, 3, 3).contiguous()
    out = _ExtractMCFn.apply(volume_rfft.contiguous(), R, tuple(image_shape), fftfreq_max, x_tolerance)
    return out.reshape(*stack_shape, *out.shape[-3:])


class _InsertFn(torch.autograd.Function):
    """(volume, weights) = scatter(slices, R); backward = adjoint gather + rotation grad."""

    @staticmethod
    def forward(ctx, slices, R, volume_shape, fftfreq_max):
        D, H, W = (int(s) for s in volume_shape)
        Wr = _rfft_width(W)
        n = R.shape[0]
        c64 = slices.dtype == torch.complex64
        vol = torch.empty((D, H, Wr), dtype=slices.dtype, device=slices.device)
        wts = torch.empty((D, H, Wr), dtype=torch.float32 if c64 else torch.float64, device=slices.device)
        _lib.call(
            "tt_insert_c64" if c64 else "tt_insert_c128", _lib.ptr(slices), D, H, W, _lib.ptr(R), n,
            _fmax_arg(fftfreq_max), _lib.ptr(vol), _lib.ptr(wts), _scatter_flags(),
            _lib.stream_of(slices),
        )
        ctx.save_for_backward(slices, R)
        ctx.geom = (D, H, W, fftfreq_max)
        return vol, wts

    @staticmethod
    def backward(ctx, grad_volume, grad_weights):
        slices, R = ctx.saved_tensors
        D, H, W, fftfreq_max = ctx.geom
        need_s, need_r = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_s or need_r):
            return None, None, None, None
        c64 = slices.dtype == torch.complex64
        if grad_volume is None:
            grad_volume = torch.zeros((D, H, _rfft_width(W)), dtype=slices.dtype, device=slices.device)
        gV = grad_volume.to(slices.dtype).contiguous()
        gW = None
        if grad_weights is not None:
            gW = grad_weights.to(torch.float32 if c64 else torch.float64).contiguous()
        gS = torch.empty_like(slices) if need_s else None
        gR = torch.zeros((R.shape[0], 3, 3), dtype=torch.float64, device=R.device) if need_r else None
        _lib.call(
            "tt_insert_bwd_c64" if c64 else "tt_insert_bwd_c128", _lib.ptr(gV), _lib.ptr(gW),
            _lib.ptr(slices), D, H, W, _lib.ptr(R), R.shape[0], _fmax_arg(fftfreq_max),
            _lib.ptr(gS), _lib.ptr(gR), _lib.stream_of(slices),
        )
        return gS, (gR.to(R.dtype) if gR is not None else None), None, None


def insert_central_slices_rfft_3d(
    image_rfft: torch.Tensor,
    volume_shape: tuple[int, int, int],
    rotation_matrices: torch.Tensor,
    fftfreq_max: float | None = None,
    zyx_matrices: bool = False,
) -> tuple[torch.Tensor, torch.Tensor]:
    """Insert fftshifted rfft slices into a 3-D rfft volume.

    Reference: slice_insertion/_insert_central_slices_rfft_3d.py:9-106.
    ``image_rfft`` is (..., h, w//2+1) complex64 or complex128; returns the
    accumulated volume (d, h, w//2+1) of the same dtype and the weights
    (float32 for complex64, float64 otherwise).
    """
    _lib.require_gpu(image_rfft)
    if image_rfft.dtype not in (torch.complex64, torch.complex128):
        raise TypeError(f"insert_central_slices_rfft_3d: complex input required, got {image_rfft.dtype}")
    D, H, W = (int(s) for s in volume_shape)
    Wr = _rfft_width(W)
    if tuple(image_rfft.shape[-2:]) != (H, Wr):
        raise IndexError(
            f"The shape of the mask [{H}, {Wr}] does not match the shape of the indexed tensor "
            f"{list(image_rfft.shape)}"
        )
    R = _prepare_rotations(rotation_matrices, image_rfft.device, zyx_matrices, None)
    if tuple(image_rfft.shape[:-2]) != tuple(R.shape[:-2]):
        raise ValueError("One coordinate triplet is required for each value in data.")
    slices = image_rfft.reshape(-1, H, Wr).contiguous()
    R = R.reshape(-1, 3, 3).contiguous()
    return _InsertFn.apply(slices, R, (D, H, W), fftfreq_max)


def insert_central_slices_rfft_3d_multichannel(
    image_rfft: torch.Tensor,
    volume_shape: tuple[int, int, int],
    rotation_matrices: torch.Tensor,
    fftfreq_max: float | None = None,
    zyx_matrices: bool = False,
) -> tuple[torch.Tensor, torch.Tensor]:
    """Multichannel insertion: (..., c, h, w//2+1) -> ((c, d, h, w//2+1), weights).

    Reference: slice_insertion/_insert_central_slices_rfft_3d.py:109-210; the
    weights are shared by the channels.  complex64: one call
    (tt_insert_mc_c64) reads every channel in place and accumulates the
    weights once; complex128 or the pixel-atomic path: a loop over channels.
    """
    _lib.require_gpu(image_rfft)
    if image_rfft.dtype not in (torch.complex64, torch.complex128):
        raise TypeError(f"insert_central_slices_rfft_3d_multichannel: complex input required, got {image_rfft.dtype}")
    if image_rfft.dtype == torch.complex128 or _config.pixel_atomic_enabled():
        vols, wts = [], None
        for c in range(image_rfft.shape[-3]):
            v, w = insert_central_slices_rfft_3d(
                image_rfft[..., c, :, :], volume_shape, rotation_matrices, fftfreq_max, zyx_matrices
            )
            vols.append(v)
            wts = w
        return torch.stack(vols, dim=0), wts
    D, H, W = (int(s) for s in volume_shape)
    Wr = _rfft_width(W)
    C = image_rfft.shape[-3]
    if tuple(image_rfft.shape[-2:]) != (H, Wr):
        raise IndexError(
            f"The shape of the mask [{H}, {Wr}] does not match the shape of the indexed tensor "
            f"{list(image_rfft.shape)}"
        )
    R = _prepare_rotations(rotation_matrices, image_rfft.device, zyx_matrices, None)
    if tuple(image_rfft.shape[:-3]) != tuple(R.shape[:-2]):
        raise ValueError("One coordinate triplet is required for each value in data.")
    slices = image_rfft.reshape(-1, C, H, Wr).contiguous()
    R = R.reshape(-1, 3, 3).contiguous()
    return _InsertMCFn.apply(slices, R, (D, H, W), fftfreq_max)


class _InsertMCFn(torch.autograd.Function):
    """(c, d, h, wr) volumes and shared weights from (n, c, h, wr) slices;
    backward: the adjoint gather of every channel, and the rotation gradient
    summed over the channels (the weights' part counted once)."""

    @staticmethod
    def forward(ctx, slices, R, volume_shape, fftfreq_max):
        D, H, W = (int(s) for s in volume_shape)
        Wr = _rfft_width(W)
        n, C = slices.shape[0], slices.shape[1]
        vol = torch.empty((C, D, H, Wr), dtype=slices.dtype, device=slices.device)
        wts = torch.empty((D, H, Wr), dtype=torch.float32, device=slices.device)
        _lib.call("tt_insert_mc_c64", _lib.ptr(slices), C, D, H, W, _lib.ptr(R), n, _fmax_arg(fftfreq_max),
                  _lib.ptr(vol), _lib.ptr(wts), _scatter_flags(), _lib.stream_of(slices))
        ctx.save_for_backward(slices, R)
        ctx.geom = (D, H, W, fftfreq_max)
        return vol, wts

    @staticmethod
    def backward(ctx, grad_volume, grad_weights):
        slices, R = ctx.saved_tensors
        D, H, W, fftfreq_max = ctx.geom
        need_s, need_r = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_s or need_r):
            return None, None, None, None
        C = slices.shape[1]
        gS = torch.empty_like(slices) if need_s else None
        gR = torch.zeros((R.shape[0], 3, 3), dtype=torch.float64, device=R.device) if need_r else None
        gW = grad_weights.to(torch.float32).contiguous() if grad_weights is not None else None
        for c in range(C):
            gV = (grad_volume[c].to(slices.dtype).contiguous() if grad_volume is not None
                  else torch.zeros((D, H, _rfft_width(W)), dtype=slices.dtype, device=slices.device))
            sc = slices[:, c].contiguous() if need_r else None
            gs = torch.empty((slices.shape[0], H, _rfft_width(W)), dtype=slices.dtype,
                             device=slices.device) if need_s else None
            _lib.call("tt_insert_bwd_c64", _lib.ptr(gV), _lib.ptr(gW if c == 0 else None), _lib.ptr(sc),
                      D, H, W, _lib.ptr(R), R.shape[0], _fmax_arg(fftfreq_max), _lib.ptr(gs), _lib.ptr(gR),
                      _lib.stream_of(slices))
            if need_s:
                gS[:, c] = gs
        return gS, (gR.to(R.dtype) if gR is not None else None), None, None


def _sinc2_grid(shape, device):
    """sinc(|f|)^2 on the fftshifted full frequency grid (TFS/project.py:58-65)."""
    axes = [torch.fft.fftshift(torch.fft.fftfreq(n, device=device)) for n in shape]
    grids = torch.meshgrid(*axes, indexing="ij")
    r = sum(g * g for g in grids) ** 0.5
    return torch.sinc(r) ** 2


class _PadSinc2Shift(torch.autograd.Function):
    """fftshift(pad(volume) * sinc^2) in one pass (TFS/project.py:44-66); the
    backward is the adjoint, crop(ifftshift(g) * sinc^2), also one pass."""

    @staticmethod
    def forward(ctx, volume, pad):
        _lib.require_gpu(volume)
        if volume.dtype not in (torch.float32, torch.float64):
            raise TypeError(f"project_3d_to_2d: volume must be float32 or float64, got {volume.dtype}")
        v = volume.contiguous()
        d = v.shape[-1]
        N = d + 2 * pad
        out = torch.empty((N, N, N), dtype=v.dtype, device=v.device)
        fn = "tt_pad_sinc2_fftshift_f32" if v.dtype == torch.float32 else "tt_pad_sinc2_fftshift_f64"
        _lib.call(fn, _lib.ptr(v), d, pad, _lib.ptr(out), _lib.stream_of(v))
        ctx.pad = pad
        return out

    @staticmethod
    def backward(ctx, grad):
        g = grad.contiguous()
        N = g.shape[-1]
        out = torch.empty((N - 2 * ctx.pad,) * 3, dtype=g.dtype, device=g.device)
        fn = "tt_ifftshift_sinc2_crop_f32" if g.dtype == torch.float32 else "tt_ifftshift_sinc2_crop_f64"
        _lib.call(fn, _lib.ptr(g), N, ctx.pad, 0, _lib.ptr(out), _lib.stream_of(g))
        return out, None


def project_3d_to_2d(
    volume: torch.Tensor,
    rotation_matrices: torch.Tensor,
    pad_factor: float = 2.0,
    fftfreq_max: float | None = None,
    zyx_matrices: bool = False,
) -> torch.Tensor:
    """Project a cubic volume by sampling central slices of its DFT.

    Reference: torch_fourier_slice/project.py:12-93 (padding, sinc^2 gridding
    correction and FFTs on torch; the central-slice gather on gfx950).
    """
    d, h, w = volume.shape[-3:]
    if len({d, h, w}) != 1:
        raise ValueError("all dimensions of volume must be equal.")
    if pad_factor < 1.0:
        raise ValueError("pad_factor must be >= 1.0")
    p = 0
    if pad_factor > 1.0:
        p = int((volume.shape[-1] * (pad_factor - 1.0)) // 2)
    volume_shape = (d + 2 * p,) * 3
    dft = _PadSinc2Shift.apply(volume, p)  # pad, x sinc^2, fftshift: one pass
    dft = torch.fft.rfftn(dft, dim=(-3, -2, -1))
    dft = torch.fft.fftshift(dft, dim=(-3, -2))
    projections = extract_central_slices_rfft_3d(
        volume_rfft=dft, image_shape=volume_shape, rotation_matrices=rotation_matrices,
        fftfreq_max=fftfreq_max, zyx_matrices=zyx_matrices,
    )
    if projections.numel() == 0:  # no poses: hipFFT refuses empty batches
        # the shape irfftn (without s) and the crop give: the last axis is
        # 2 (Wr - 1), one less than the padded box when that is odd
        hp, wr = projections.shape[-2:]
        shape = (hp - 2 * p, 2 * (wr - 1) - 2 * p)
        return torch.empty(projections.shape[:-2] + shape, dtype=volume.dtype, device=volume.device)
    projections = torch.fft.ifftshift(projections, dim=(-2,))
    projections = torch.fft.irfftn(projections, dim=(-2, -1))
    projections = torch.fft.ifftshift(projections, dim=(-2, -1))
    if pad_factor > 1.0:
        projections = F.pad(projections, pad=[-p] * 4)
    return projections


def project_3d_to_2d_multichannel(
    volume: torch.Tensor,
    rotation_matrices: torch.Tensor,
    pad_factor: float = 2.0,
    fftfreq_max: float | None = None,
    zyx_matrices: bool = False,
) -> torch.Tensor:
    """Multichannel projection (c, d, d, d) -> (..., c, d, d) (TFS/project.py:96-177).

    The fused pad / sinc^2 / fftshift prologue per channel, one batched rfftn,
    then one multichannel extraction (the brick gather serves two channels
    per launch), as the reference makes one multichannel extraction call.
    """
    d, h, w = volume.shape[-3:]
    if len({d, h, w}) != 1:
        raise ValueError("all dimensions of volume must be equal.")
    if pad_factor < 1.0:
        raise ValueError("pad_factor must be >= 1.0")
    p = 0
    if pad_factor > 1.0:
        p = int((volume.shape[-1] * (pad_factor - 1.0)) // 2)
    volume_shape = (d + 2 * p,) * 3
    dft = torch.stack([_PadSinc2Shift.apply(volume[c], p) for c in range(volume.shape[0])], dim=0)
    dft = torch.fft.rfftn(dft, dim=(-3, -2, -1))
    dft = torch.fft.fftshift(dft, dim=(-3, -2))
    projections = extract_central_slices_rfft_3d_multichannel(
        volume_rfft=dft, image_shape=volume_shape, rotation_matrices=rotation_matrices,
        fftfreq_max=fftfreq_max, zyx_matrices=zyx_matrices,
    )
    if projections.numel() == 0:  # no poses: hipFFT refuses empty batches
        # the shape irfftn (without s) and the crop give: the last axis is
        # 2 (Wr - 1), one less than the padded box when that is odd
        hp, wr = projections.shape[-2:]
        shape = (hp - 2 * p, 2 * (wr - 1) - 2 * p)
        return torch.empty(projections.shape[:-2] + shape, dtype=volume.dtype, device=volume.device)
    projections = torch.fft.ifftshift(projections, dim=(-2,))
    projections = torch.fft.irfftn(projections, dim=(-2, -1))
    projections = torch.fft.ifftshift(projections, dim=(-2, -1))
    if pad_factor > 1.0:
        projections = F.pad(projections, pad=[-p] * 4)
    return projections


def backproject_2d_to_3d(
    images: torch.Tensor,
    rotation_matrices: torch.Tensor,
    pad_factor: float = 2.0,
    fftfreq_max: float | None = None,
    zyx_matrices: bool = False,
) -> torch.Tensor:
    """Reconstruct a cubic volume from 2-D projections by central-slice insertion.

    Reference: torch_fourier_slice/backproject.py:11-92.
    """
    h, w = images.shape[-2:]
    if h != w:
        raise ValueError("images must be square.")
    if pad_factor < 1.0:
        raise ValueError("pad_factor must be >= 1.0")
    p = 0
    if pad_factor > 1.0:
        p = int((images.shape[-1] * (pad_factor - 1.0)) // 2)
        images = F.pad(images, pad=[p] * 4)
    h, w = images.shape[-2:]
    volume_shape = (w, w, w)
    if images.numel() == 0:  # hipFFT refuses empty batches: an empty rfft stack
        cdt = torch.complex128 if images.dtype == torch.float64 else torch.complex64
        images = torch.empty(images.shape[:-1] + (w // 2 + 1,), dtype=cdt, device=images.device)
    else:
        images = torch.fft.fftshift(images, dim=(-2, -1))
        images = torch.fft.rfftn(images, dim=(-2, -1))
        images = torch.fft.fftshift(images, dim=(-2,))
    dft, weights = insert_central_slices_rfft_3d(
        image_rfft=images, volume_shape=volume_shape, rotation_matrices=rotation_matrices,
        fftfreq_max=fftfreq_max, zyx_matrices=zyx_matrices,
    )
    if torch.is_grad_enabled() and (dft.requires_grad or weights.requires_grad):
        # differentiable path: the same steps as torch operations
        weights = torch.clamp(weights, min=1.0)
        dft = dft / weights
        dft = torch.fft.ifftshift(dft, dim=(-3, -2))
        dft = torch.fft.irfftn(dft, dim=(-3, -2, -1))
        dft = torch.fft.ifftshift(dft, dim=(-3, -2, -1))
        dft = dft / _sinc2_grid(dft.shape[-3:], dft.device)
        if pad_factor > 1.0:
            dft = F.pad(dft, pad=[-p] * 6)
        return torch.real(dft)
    # fused: divide by clamp(weights, 1) with the ifftshift, then ifftshift,
    # / sinc^2 and the crop after the inverse FFT, one pass each
    D, H, Wr = dft.shape
    shifted = torch.empty_like(dft)
    fn = "tt_wdiv_ifftshift_c64" if dft.dtype == torch.complex64 else "tt_wdiv_ifftshift_c128"
    _lib.call(fn, _lib.ptr(dft), _lib.ptr(weights.contiguous()), D, H, Wr, _lib.ptr(shifted), _lib.stream_of(dft))
    real = torch.fft.irfftn(shifted, dim=(-3, -2, -1)).contiguous()
    N = real.shape[-1]
    out = torch.empty((N - 2 * p,) * 3, dtype=real.dtype, device=real.device)
    fn = "tt_ifftshift_sinc2_crop_f32" if real.dtype == torch.float32 else "tt_ifftshift_sinc2_crop_f64"
    _lib.call(fn, _lib.ptr(real), N, p, 1, _lib.ptr(out), _lib.stream_of(real))
    return out


def backproject_2d_to_3d_multichannel(
    images: torch.Tensor,
    rotation_matrices: torch.Tensor,
    pad_factor: float = 2.0,
    fftfreq_max: float | None = None,
    zyx_matrices: bool = False,
) -> torch.Tensor:
    """Multichannel reconstruction (..., c, d, d) -> (c, d, d, d) (TFS/backproject.py:95-181)."""
    outs = [
        backproject_2d_to_3d(images[..., c, :, :], rotation_matrices, pad_factor, fftfreq_max, zyx_matrices)
        for c in range(images.shape[-3])
    ]
    return torch.stack(outs, dim=0)


# ---------------------------------------------------------------- 2-D analogues


class _ExtractLineFn(torch.autograd.Function):
    """lines = gather(image, R_yx); backward = adjoint splat + rotation grad."""

    @staticmethod
    def forward(ctx, image, R, image_shape, fftfreq_max):
        n = R.shape[0]
        h, wr = image.shape
        img_h, img_w = (int(v) for v in image_shape)
        out = torch.empty((n, _rfft_width(img_w)), dtype=image.dtype, device=image.device)
        _lib.call("tt_extract2d_c64", _lib.ptr(image), h, wr, img_h, img_w, _lib.ptr(R), n,
                  _fmax_arg(fftfreq_max), _lib.ptr(out), _lib.stream_of(image))
        ctx.save_for_backward(image, R)
        ctx.geom = (img_h, img_w, fftfreq_max)
        return out

    @staticmethod
    def backward(ctx, grad_lines):
        image, R = ctx.saved_tensors
        img_h, img_w, fftfreq_max = ctx.geom
        need_i, need_r = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_i or need_r):
            return None, None, None, None
        g = grad_lines.to(torch.complex64).contiguous()
        h, wr = image.shape
        gI = torch.zeros_like(image) if need_i else None
        gR = torch.zeros((R.shape[0], 2, 2), dtype=torch.float64, device=R.device) if need_r else None
        _lib.call("tt_extract2d_bwd_c64", _lib.ptr(g), _lib.ptr(image), h, wr, img_h, img_w, _lib.ptr(R),
                  R.shape[0], _fmax_arg(fftfreq_max), _lib.ptr(gI), _lib.ptr(gR), _lib.stream_of(image))
        return gI, (gR.to(R.dtype) if gR is not None else None), None, None


def extract_central_slices_rfft_2d(
    image_rfft: torch.Tensor,
    image_shape: tuple[int, int],
    rotation_matrices: torch.Tensor,
    fftfreq_max: float | None = None,
    yx_matrices: bool = False,
) -> torch.Tensor:
    """Extract central lines from an fftshifted 2-D rfft.

    Reference: slice_extraction/_extract_central_slices_rfft_2d.py:9-92.
    ``image_rfft`` is (h, w//2+1) complex64 on a ROCm device; returns
    (..., w//2+1) for rotation matrices of shape (..., 2, 2).  The reference
    raises einops.EinopsError for every ``fftfreq_max`` (its elementwise mask,
    :34-36, flattens the grid); here the line is cut at |f| <= fftfreq_max.
    """
    _lib.require_gpu(image_rfft)
    if image_rfft.ndim != 2:
        raise ValueError(f"image_rfft should have shape (h, w//2+1), got {tuple(image_rfft.shape)}")
    if image_rfft.dtype != torch.complex64:
        raise RuntimeError(
            "extract_central_slices_rfft_2d: image_rfft must be complex64 "
            f"(grid_sample needs input and grid of the same dtype), got {image_rfft.dtype}"
        )
    R = rotation_matrices.to(device=image_rfft.device, dtype=torch.float32)
    if not yx_matrices:
        R = torch.flip(R, dims=(-2, -1))
    stack_shape = tuple(R.shape[:-2])
    R = R.reshape(-1, 2, 2).contiguous()
    out = _ExtractLineFn.apply(image_rfft.contiguous(), R, tuple(image_shape), fftfreq_max)
    return out.reshape(*stack_shape, out.shape[-1])


def _sinc2_grid_2d(shape, device):
    """sinc(|f|)^2 on the fftshifted full 2-D frequency grid (TFS/project.py:218-225)."""
    axes = [torch.fft.fftshift(torch.fft.fftfreq(n, device=device)) for n in shape]
    gy, gx = torch.meshgrid(*axes, indexing="ij")
    return torch.sinc((gy * gy + gx * gx) ** 0.5) ** 2


def project_2d_to_1d(
    image: torch.Tensor,
    rotation_matrices: torch.Tensor,
    pad_factor: float = 2.0,
    fftfreq_max: float | None = None,
    yx_matrices: bool = False,
) -> torch.Tensor:
    """Project a square image by sampling central lines of its DFT.

    Reference: torch_fourier_slice/project.py:180-260 (padding, sinc^2
    gridding correction and FFTs on torch; the central-line gather on gfx950).
    Returns (..., d) for rotation matrices of shape (..., 2, 2).
    """
    h, w = image.shape[-2:]
    if h != w:
        raise ValueError("all dimensions of image must be equal.")
    if pad_factor < 1.0:
        raise ValueError("pad_factor must be >= 1.0")
    p = 0
    if pad_factor > 1.0:
        p = int((image.shape[-1] * (pad_factor - 1.0)) // 2)
        image = F.pad(image, pad=[p] * 4)
    image_shape = tuple(image.shape[-2:])
    image = image * _sinc2_grid_2d(image_shape, image.device)
    dft = torch.fft.fftshift(image, dim=(-2, -1))
    dft = torch.fft.rfftn(dft, dim=(-2, -1))
    dft = torch.fft.fftshift(dft, dim=(-2,))
    lines = extract_central_slices_rfft_2d(
        image_rfft=dft, image_shape=image_shape, rotation_matrices=rotation_matrices,
        fftfreq_max=fftfreq_max, yx_matrices=yx_matrices,
    )
    if lines.numel() == 0:  # no rotations: hipFFT refuses empty batches
        n = 2 * (lines.shape[-1] - 1) - 2 * p  # irfftn's length (without s), then the crop
        return torch.empty(lines.shape[:-1] + (n,), dtype=image.dtype, device=image.device)
    projections = torch.fft.irfftn(lines, dim=(-1,))
    projections = torch.fft.ifftshift(projections, dim=(-1,))
    if pad_factor > 1.0:
        projections = F.pad(projections, pad=[-p] * 2)
    return projections
