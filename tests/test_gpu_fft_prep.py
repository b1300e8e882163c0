"""The fused elementwise steps around the FFTs of project_3d_to_2d and
backproject_2d_to_3d against the torch operations they replace
(TFS/project.py:44-70, TFS/backproject.py:75-91), their adjointness, and the
fused backprojection against its differentiable torch path."""

import pytest
import torch
import torch.nn.functional as F

import teamtomo_amd as tt
from teamtomo_amd import _lib
from teamtomo_amd.fourier_slice import _PadSinc2Shift, _sinc2_grid
from teamtomo_amd.synthetic import uniform_rotations

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("d,p,dtype", [(16, 8, torch.float32), (15, 7, torch.float64), (12, 0, torch.float32),
                                       (9, 4, torch.float32)])
def test_pad_sinc2_shift_matches_torch_and_is_adjoint(device, d, p, dtype):
    g = torch.Generator(device=device).manual_seed(d)
    v = torch.randn((d, d, d), generator=g, device=device, dtype=dtype)
    out = _PadSinc2Shift.apply(v, p)
    N = d + 2 * p
    ref = torch.fft.fftshift(F.pad(v, [p] * 6) * _sinc2_grid((N, N, N), device).to(dtype), dim=(-3, -2, -1))
    # the kernel evaluates sinc^2 in FP64, torch's grid is float32 (as in the
    # reference): equal to float32 rounding of the grid for both dtypes
    assert float((out - ref).abs().max()) <= 1e-6 * float(ref.abs().max())
    # backward = adjoint: <A v, u> = <v, A^T u>
    u = torch.randn((N, N, N), generator=g, device=device, dtype=dtype)
    vv = v.clone().requires_grad_(True)
    torch.sum(_PadSinc2Shift.apply(vv, p) * u).backward()
    lhs = float(torch.sum(out.double() * u.double()))
    rhs = float(torch.sum(v.double() * vv.grad.double()))
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("pad_factor,dtype", [(2.0, torch.float32), (1.0, torch.float32), (2.0, torch.float64)])
def test_fused_backprojection_equals_torch_path(device, pad_factor, dtype):
    g = torch.Generator(device=device).manual_seed(7)
    imgs = torch.randn((9, 20, 20), generator=g, device=device, dtype=dtype)
    R = uniform_rotations(9, 8, device=device)
    fused = tt.backproject_2d_to_3d(imgs, R, pad_factor=pad_factor)
    ref = tt.backproject_2d_to_3d(imgs.clone().requires_grad_(True), R, pad_factor=pad_factor).detach()
    assert fused.shape == ref.shape == (20, 20, 20) and fused.dtype == dtype
    assert float((fused - ref).abs().max()) <= 1e-5 * float(ref.abs().max())


def test_fused_prologue_epilogue_timing(device):
    # 256^3 padded volume: the fused passes against the torch operations they replace
    import time

    g = torch.Generator(device=device).manual_seed(9)
    v = torch.randn((128, 128, 128), generator=g, device=device)
    dft = torch.complex(torch.randn((256, 256, 129), generator=g, device=device),
                        torch.randn((256, 256, 129), generator=g, device=device))
    w = torch.rand((256, 256, 129), generator=g, device=device) * 3
    N, p = 256, 64

    def fused():
        _PadSinc2Shift.apply(v, p)
        out = torch.empty_like(dft)
        _lib.call("tt_wdiv_ifftshift_c64", _lib.ptr(dft), _lib.ptr(w), 256, 256, 129, _lib.ptr(out), _lib.stream_of(dft))
        real = torch.randn((N, N, N), device=device)
        res = torch.empty((128, 128, 128), device=device)
        _lib.call("tt_ifftshift_sinc2_crop_f32", _lib.ptr(real), N, p, 1, _lib.ptr(res), _lib.stream_of(real))

    def torch_ops():
        torch.fft.fftshift(F.pad(v, [p] * 6) * _sinc2_grid((N, N, N), device), dim=(-3, -2, -1))
        torch.fft.ifftshift(dft / torch.clamp(w, min=1.0), dim=(-3, -2))
        real = torch.randn((N, N, N), device=device)
        F.pad(torch.fft.ifftshift(real, dim=(-3, -2, -1)) / _sinc2_grid((N, N, N), device), [-p] * 6)

    t = {}
    for name, f in (("fused", fused), ("torch", torch_ops)):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(10):
            f()
        torch.cuda.synchronize()
        t[name] = (time.perf_counter() - t0) / 10 * 1e3
    print(f"project/backproject elementwise steps at 256^3 (incl. one 64 MB randn each): fused {t['fused']:.2f} ms, "
          f"torch {t['torch']:.2f} ms")
    # one pass per step against several torch passes and temporaries
    assert t["fused"] < t["torch"]
