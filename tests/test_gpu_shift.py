"""GPU parity of the Fourier shifts (torch-fourier-shift drop-ins) against the
FP64 oracle and the reference's own test cases, through the C ABI
(tt_fourier_shift_*).

Tolerances: complex64 results within 2e-6 relative L2 of the FP64 oracle
(the phase is reduced in FP64, so large shifts do not lose accuracy);
complex128 within 1e-12; gradients within 1e-5 (c64) of the oracle's
analytic gradient, itself pinned against finite differences.
"""

import time

import numpy as np
import pytest
import torch

import teamtomo_amd as tt
from conftest import rel_l2
from oracle import fourier_shift as O

pytestmark = pytest.mark.gpu


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def _np(t):
    return t.detach().cpu().numpy()


FN = {1: tt.fourier_shift_dft_1d, 2: tt.fourier_shift_dft_2d, 3: tt.fourier_shift_dft_3d}
IMG = {1: tt.fourier_shift_image_1d, 2: tt.fourier_shift_image_2d, 3: tt.fourier_shift_image_3d}
SHAPES = {1: (30,), 2: (24, 31), 3: (10, 12, 9)}


def _stored(shape, rfft):
    return shape[:-1] + (shape[-1] // 2 + 1,) if rfft else shape


def _cplx(rng, shape, dtype=np.complex64):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)


@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("rfft", [False, True])
@pytest.mark.parametrize("fftshifted", [False, True])
def test_shift_dft_vs_oracle(device, nd, rfft, fftshifted):
    rng = np.random.default_rng(100 * nd + 10 * rfft + fftshifted)
    shape = SHAPES[nd]
    st = _stored(shape, rfft)
    dft = _cplx(rng, (4,) + st)
    s = rng.uniform(-40, 40, (4, nd)).astype(np.float32)
    s_arg = s[:, 0] if nd == 1 else s
    out = FN[nd](_t(dft, device), shape, _t(s_arg, device), rfft, fftshifted)
    ref = O.fourier_shift_dft(dft, shape, s_arg, rfft, fftshifted)
    assert out.shape == ref.shape and out.dtype == torch.complex64
    assert rel_l2(_np(out), ref) < 2e-6
    # complex128 with float64 shifts: the reference's promotion, FP64 throughout
    out64 = FN[nd](_t(dft.astype(np.complex128), device), shape, _t(s_arg.astype(np.float64), device), rfft,
                   fftshifted)
    assert out64.dtype == torch.complex128
    assert rel_l2(_np(out64), O.fourier_shift_dft(dft.astype(np.complex128), shape, s_arg, rfft, fftshifted)) < 1e-12


def test_shift_broadcasting_and_dtypes(device):
    rng = np.random.default_rng(7)
    shape = (16, 20)
    st = _stored(shape, True)
    one = _cplx(rng, st)
    s = rng.uniform(-5, 5, (6, 2)).astype(np.float32)
    # one DFT, many shifts (served without a copy)
    out = tt.fourier_shift_dft_2d(_t(one, device), shape, _t(s, device), True, True)
    assert out.shape == (6,) + st
    assert rel_l2(_np(out), O.fourier_shift_dft(one, shape, s, True, True)) < 2e-6
    # (4, 1) DFTs against (1, 3) shifts -> (4, 3)
    many = _cplx(rng, (4, 1) + st)
    s13 = rng.uniform(-5, 5, (1, 3, 2)).astype(np.float32)
    out = tt.fourier_shift_dft_2d(_t(many, device), shape, _t(s13, device), True, False)
    assert out.shape == (4, 3) + st
    assert rel_l2(_np(out), O.fourier_shift_dft(many, shape, s13, True, False)) < 2e-6
    # integer shifts (the reference's own tests pass them) and float64 shifts
    si = np.array([[1, 2]])
    out = tt.fourier_shift_dft_2d(_t(one, device), shape, _t(si, device), True, False)
    assert out.dtype == torch.complex64
    assert rel_l2(_np(out), O.fourier_shift_dft(one, shape, si, True, False)) < 2e-6
    out = tt.fourier_shift_dft_2d(_t(one, device), shape, _t(s.astype(np.float64), device), True, False)
    assert out.dtype == torch.complex128
    # a real "DFT" is promoted like the reference's product
    out = tt.fourier_shift_dft_1d(torch.ones(8, device=device), (8,), torch.tensor(0.5, device=device), False, False)
    assert out.dtype == torch.complex64 and out.shape == (8,)
    # mismatched grids raise
    with pytest.raises(RuntimeError):
        tt.fourier_shift_dft_2d(_t(one, device), (16, 22), _t(s, device), True, False)
    with pytest.raises(ValueError):
        tt.fourier_shift_dft_2d(_t(one, device), shape, _t(s[:, :1], device), True, False)
    with pytest.raises(RuntimeError):
        tt.fourier_shift_dft_2d(_t(one, "cpu"), shape, _t(s, "cpu"), True, False)


def test_large_shifts_keep_precision(device):
    """Phases reduced in FP64: shifts of thousands of pixels stay accurate in
    complex64 (a float32 angle of up to 2 pi x 2500 cycles rounds in steps of ~1e-3 rad)."""
    rng = np.random.default_rng(3)
    shape = (64, 64)
    dft = _cplx(rng, (8,) + _stored(shape, True))
    s = rng.uniform(-5000, 5000, (8, 2)).astype(np.float32)
    out = tt.fourier_shift_dft_2d(_t(dft, device), shape, _t(s, device), True, True)
    assert rel_l2(_np(out), O.fourier_shift_dft(dft, shape, s, True, True)) < 2e-6


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_reference_delta_images(device, nd):
    """tests/test_phase_shift_{1,2,3}d.py: a delta at index 2 of a 4^nd image
    moves by +1 along each shifted axis (-1 too in 1-D)."""
    img = torch.zeros((4,) * nd, device=device)
    img[(2,) * nd] = 1
    cases = [np.ones(nd)] + [np.eye(nd)[a] for a in range(nd)] if nd > 1 else [np.ones(1), -np.ones(1)]
    for s in cases:
        shifts = torch.tensor(s[None] if nd > 1 else s, dtype=torch.float32, device=device)
        out = IMG[nd](img, shifts)
        want = torch.zeros_like(img)
        want[tuple(int(2 + v) for v in s)] = 1
        assert torch.allclose(out[0], want, atol=1e-5), (nd, s)


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_shift_image_vs_oracle(device, nd):
    rng = np.random.default_rng(50 + nd)
    shape = {1: (64,), 2: (40, 33), 3: (12, 16, 20)}[nd]
    img = rng.standard_normal((5,) + shape).astype(np.float32)
    s = rng.uniform(-6, 6, (5, nd)).astype(np.float32)
    s_arg = s[:, 0] if nd == 1 else s
    out = IMG[nd](_t(img, device), _t(s_arg, device))
    ref = O.fourier_shift_image_nd(img, s_arg, nd)
    assert out.shape == ref.shape  # odd last axes come back one shorter, as in the reference
    assert rel_l2(_np(out), ref) < 2e-6


@pytest.mark.parametrize("rfft,fftshifted", [(True, True), (False, False)])
def test_shift_gradients_vs_oracle(device, rfft, fftshifted):
    rng = np.random.default_rng(21)
    shape = (24, 30)
    st = _stored(shape, rfft)
    dft = _cplx(rng, (3,) + st)
    g = _cplx(rng, (3,) + st)
    s = rng.uniform(-4, 4, (3, 2)).astype(np.float32)
    D = _t(dft, device).requires_grad_(True)
    S = _t(s, device).requires_grad_(True)
    out = tt.fourier_shift_dft_2d(D, shape, S, rfft, fftshifted)
    torch.real(torch.sum(torch.conj(_t(g, device)) * out)).backward()
    ref_gs = O.shift_grad(g.astype(np.complex128), dft.astype(np.complex128), s, shape, rfft, fftshifted)
    assert rel_l2(_np(S.grad), ref_gs) < 1e-5
    ref_gd = g * np.conj(O.phase_shift_grid(s, shape, rfft, fftshifted))
    assert rel_l2(_np(D.grad), ref_gd) < 2e-6
    # one DFT shared by every shift: its gradient sums over the shifts
    D1 = _t(dft[0], device).requires_grad_(True)
    out = tt.fourier_shift_dft_2d(D1, shape, _t(s, device), rfft, fftshifted)
    torch.real(torch.sum(torch.conj(_t(g, device)) * out)).backward()
    ref = np.sum(g * np.conj(O.phase_shift_grid(s, shape, rfft, fftshifted)), axis=0)
    assert rel_l2(_np(D1.grad), ref) < 2e-6


def test_shift_gradient_complex128_vs_finite_differences(device):
    rng = np.random.default_rng(31)
    shape = (8, 10, 12)
    st = _stored(shape, True)
    dft = _t(_cplx(rng, (2,) + st, np.complex128), device)
    g = _t(_cplx(rng, (2,) + st, np.complex128), device)
    s = torch.tensor(rng.uniform(-3, 3, (2, 3)), device=device, requires_grad=True)

    def loss(sh):
        return torch.real(torch.sum(torch.conj(g) * tt.fourier_shift_dft_3d(dft, shape, sh, True, False)))

    loss(s).backward()
    h = 1e-6
    with torch.no_grad():
        for b in range(2):
            for a in range(3):
                e = torch.zeros_like(s)
                e[b, a] = h
                fd = (loss(s + e) - loss(s - e)).item() / (2 * h)
                assert abs(s.grad[b, a].item() - fd) < 1e-6 * max(1.0, abs(fd)), (b, a)


def test_shift_pass_rate(device):
    """A 61-tilt batch of 512^2 rffts (fftshifted): one pass of
    the kernel against the reference's composition in torch (fftfreq grid,
    angle sum, cos/sin, complex, fftshift, product) on the same GPU."""
    n, h, w = 61, 512, 512
    wr = w // 2 + 1
    dft = torch.randn(n, h, wr, dtype=torch.complex64, device=device)
    s = (torch.rand(n, 2, device=device) - 0.5) * 10

    def ours():
        return tt.fourier_shift_dft_2d(dft, (h, w), s, True, True)

    def torch_ref():
        fy = torch.fft.fftfreq(h, device=device)
        fx = torch.fft.rfftfreq(w, device=device)
        grid = torch.stack(torch.meshgrid(fy, fx, indexing="ij"), -1)
        ang = torch.sum(-2 * torch.pi * grid * s[:, None, None, :], -1)
        ph = torch.fft.fftshift(torch.complex(torch.cos(ang), torch.sin(ang)), dim=(-2,))
        return dft * ph

    assert rel_l2(_np(ours()), _np(torch_ref())) < 1e-4

    def timed(fn, reps=20):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    t_ours, t_ref = timed(ours), timed(torch_ref)
    gbs = 2 * dft.numel() * 8 / t_ours / 1e9
    print(f"\nfourier_shift_dft_2d 61x512x257 c64: {t_ours * 1e6:.1f} us ({gbs:.0f} GB/s), "
          f"torch composition {t_ref * 1e6:.1f} us ({t_ref / t_ours:.1f}x)")
    assert t_ours < t_ref


def test_empty_inputs(device):
    """Empty batches give empty results of the right shape and dtype."""
    dft = torch.zeros((0, 8, 5), dtype=torch.complex64, device=device)
    out = tt.fourier_shift_dft_2d(dft, (8, 8), torch.zeros((0, 2), device=device), True, False)
    assert out.shape == (0, 8, 5) and out.dtype == torch.complex64
    out = tt.fourier_shift_image_3d(torch.zeros((0, 4, 4, 4), device=device), torch.zeros((0, 3), device=device))
    assert out.shape == (0, 4, 4, 4)
    assert tt.sample_image_1d(torch.rand(10, device=device), torch.zeros(0, device=device)).shape == (0,)
    img = torch.zeros(10, device=device)
    img, w = tt.insert_into_image_1d(torch.zeros(0, device=device), torch.zeros(0, device=device), img)
    assert not bool(img.abs().amax()) and not bool(w.abs().amax())
