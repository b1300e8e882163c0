"""GPU parity of the 2-D analogues (central lines, project_2d_to_1d, the 2-D
sampler and splat) against the FP64 oracle and the reference's golden
fixtures, plus their hand-written backward against the oracle's adjoint and
FP64 finite differences."""

import numpy as np
import pytest
import torch

import teamtomo_amd as tt
from conftest import golden, max_abs, rel_l2
from oracle import slice_2d as O

pytestmark = pytest.mark.gpu

EXT_TOL = 1e-5
INS_TOL = 1e-5


def _t(a, device, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(device)
    return t.to(dtype) if dtype is not None else t


def _np(t):
    return t.detach().cpu().numpy()


def _rots(n, seed):
    th = np.random.default_rng(seed).uniform(0, 2 * np.pi, n)
    c, s = np.cos(th), np.sin(th)
    return np.stack([np.stack([c, -s], -1), np.stack([s, c], -1)], -2)


@pytest.mark.parametrize("key,shape", [("even", (32, 32)), ("odd", (33, 33)), ("shape2x", (32, 32))])
def test_lines_vs_golden_and_oracle(device, key, shape):
    G = golden("extract_2d")
    F = _t(G[f"{key}_image_rfft"], device)
    R = _t(G[f"{key}_rotations"], device)
    L = _np(tt.extract_central_slices_rfft_2d(F, shape, R))
    Ly = _np(tt.extract_central_slices_rfft_2d(F, shape, R, yx_matrices=True))
    ref = O.extract_central_lines(G[f"{key}_image_rfft"], shape, O.prepare_rotations_2d(G[f"{key}_rotations"]))
    print(f"lines {key}: rel_l2 vs FP64 {rel_l2(L, ref):.2e}, vs golden {rel_l2(L, G[f'{key}_lines']):.2e}")
    assert rel_l2(L, ref) < EXT_TOL
    assert rel_l2(L, G[f"{key}_lines"]) < EXT_TOL
    assert rel_l2(Ly, G[f"{key}_lines_yx"]) < EXT_TOL
    assert np.array_equal(L == 0, ref == 0)


def test_lines_leading_dims_identity_and_fftfreq_max(device):
    rng = np.random.default_rng(7)
    Fn = rng.standard_normal((64, 33)) + 1j * rng.standard_normal((64, 33))
    F = _t(Fn, device, torch.complex64)
    R = _rots(12, 8).reshape(3, 4, 2, 2)
    L = tt.extract_central_slices_rfft_2d(F, (64, 64), _t(R, device))
    assert L.shape == (3, 4, 33)
    ref = O.extract_central_lines(Fn, (64, 64), O.prepare_rotations_2d(R))
    assert rel_l2(_np(L).reshape(12, 33), ref) < EXT_TOL
    I = tt.extract_central_slices_rfft_2d(F, (64, 64), torch.eye(2, device=device)[None])
    assert torch.equal(I[0], F[32])
    Lm = _np(tt.extract_central_slices_rfft_2d(F, (64, 64), _t(R, device), fftfreq_max=0.3))
    refm = O.extract_central_lines(Fn, (64, 64), O.prepare_rotations_2d(R), fftfreq_max=0.3)
    assert rel_l2(Lm.reshape(12, 33), refm) < EXT_TOL
    assert np.all(Lm[..., 20:] == 0)  # 20/64 > 0.3


def test_lines_large(device):
    # many rotations of a 512 x 257 rfft: parity on a subset, sizes the kernel
    # walks with several passes per workgroup
    rng = np.random.default_rng(9)
    Fn = (rng.standard_normal((512, 257)) + 1j * rng.standard_normal((512, 257))).astype(np.complex64)
    R = _rots(5000, 10)
    L = _np(tt.extract_central_slices_rfft_2d(_t(Fn, device), (512, 512), _t(R, device)))
    sub = np.arange(0, 5000, 97)
    ref = O.extract_central_lines(Fn, (512, 512), O.prepare_rotations_2d(R[sub]))
    assert rel_l2(L[sub], ref) < EXT_TOL


def test_lines_adjoint_large(device):
    # the tiled image gradient (split tiles near DC, shared tile edges) against
    # the oracle's adjoint, the dot-product identity with the forward, and
    # bitwise reruns
    rng = np.random.default_rng(15)
    Fn = (rng.standard_normal((512, 257)) + 1j * rng.standard_normal((512, 257))).astype(np.complex64)
    R = _rots(3000, 16)
    Gn = (rng.standard_normal((3000, 257)) + 1j * rng.standard_normal((3000, 257))).astype(np.complex64)
    F = _t(Fn, device).requires_grad_(True)
    L = tt.extract_central_slices_rfft_2d(F, (512, 512), _t(R, device))
    torch.real(torch.sum(torch.conj(_t(Gn, device)) * L)).backward()
    ref = O.extract_central_lines_adjoint(Gn, Fn.shape, (512, 512), O.prepare_rotations_2d(R))
    print("line adjoint 512^2 x 3000 rel_l2:", rel_l2(_np(F.grad), ref))
    assert rel_l2(_np(F.grad), ref) < 1e-5
    lhs = np.real(np.vdot(Gn.astype(np.complex128), _np(L).astype(np.complex128)))
    rhs = np.real(np.vdot(_np(F.grad).astype(np.complex128), Fn.astype(np.complex128)))
    assert abs(lhs - rhs) < 1e-4 * abs(lhs)
    # integer sums at the shared tile edges and split tiles: bitwise reruns
    for _ in range(2):
        F2 = _t(Fn, device).requires_grad_(True)
        L2 = tt.extract_central_slices_rfft_2d(F2, (512, 512), _t(R, device))
        torch.real(torch.sum(torch.conj(_t(Gn, device)) * L2)).backward()
        assert torch.equal(F2.grad, F.grad)


def test_project_2d_to_1d_vs_golden(device):
    G = golden("extract_2d")
    img = _t(G["proj_image"], device)
    R = _t(G["proj_rotations"], device)
    P2 = _np(tt.project_2d_to_1d(img, R))
    P1 = _np(tt.project_2d_to_1d(img, R, pad_factor=1.0))
    assert P2.shape == G["proj_pad2"].shape
    assert rel_l2(P2, G["proj_pad2"]) < EXT_TOL and rel_l2(P1, G["proj_pad1"]) < EXT_TOL


def test_project_2d_rotation_centre(device):
    # TFS tests/test_torch_fourier_slice.py:53-76
    img = torch.zeros((32, 32), device=device)
    img[16, 16] = 1
    P = tt.project_2d_to_1d(img, _t(_rots(100, 11), device))
    assert torch.all(torch.argmax(P, dim=-1) == 16)


def test_project_2d_errors(device):
    with pytest.raises(ValueError):
        tt.project_2d_to_1d(torch.zeros((8, 9), device=device), torch.eye(2, device=device))
    with pytest.raises(ValueError):
        tt.project_2d_to_1d(torch.zeros((8, 8), device=device), torch.eye(2, device=device), pad_factor=0.5)


def test_lines_backward_image_and_rotation(device):
    rng = np.random.default_rng(12)
    shape = (24, 13)
    Fn = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    R64 = _rots(6, 13)
    T = rng.standard_normal((6, 13)) + 1j * rng.standard_normal((6, 13))
    F = _t(Fn, device, torch.complex64).requires_grad_(True)
    R = _t(R64, device, torch.float64).requires_grad_(True)
    L = tt.extract_central_slices_rfft_2d(F, (24, 24), R)
    loss = torch.real(torch.sum(torch.conj(_t(T, device, torch.complex64)) * L))
    loss.backward()
    # image gradient: the adjoint splat of the target lines
    gI = O.extract_central_lines_adjoint(T, shape, (24, 24), O.prepare_rotations_2d(R64))
    print("line image grad rel_l2:", rel_l2(_np(F.grad), gI))
    assert rel_l2(_np(F.grad), gI) < 1e-5
    # rotation gradient: FP64 central differences of the oracle's loss
    def f(Rx):
        Lx = O.extract_central_lines(Fn, (24, 24), Rx[..., ::-1, ::-1])
        return np.real(np.sum(np.conj(T) * Lx))
    h = 1e-7
    fd = np.zeros_like(R64)
    for idx in np.ndindex(R64.shape):
        Rp, Rm = R64.copy(), R64.copy()
        Rp[idx] += h
        Rm[idx] -= h
        fd[idx] = (f(Rp) - f(Rm)) / (2 * h)
    print("line rotation grad rel_l2 vs FD:", rel_l2(_np(R.grad), fd))
    assert rel_l2(_np(R.grad), fd) < 1e-4
    assert np.all(_np(R.grad)[:, :, 1] == 0)  # the y column of R_xy meets a zero y frequency


@pytest.mark.parametrize("mode", ["bilinear", "nearest", "bicubic"])
def test_sample_image_2d_vs_golden(device, mode):
    G = golden("sample_image_2d")
    crd = _t(G["coordinates"], device)
    for key in ("real", "complex", "multi"):
        out = _np(tt.sample_image_2d(_t(G[f"image_{key}"], device), crd, interpolation=mode))
        ref = O.sample_image_2d(G[f"image_{key}"], G["coordinates"], mode)
        assert out.shape == G[f"{key}_{mode}"].shape
        assert rel_l2(out, G[f"{key}_{mode}"]) < EXT_TOL and rel_l2(out, ref) < EXT_TOL
        if mode == "nearest":
            assert max_abs(out, G[f"{key}_{mode}"]) == 0.0


def test_sample_image_2d_float64_and_errors(device):
    rng = np.random.default_rng(14)
    img = rng.standard_normal((3, 9, 7))
    c = np.stack([rng.uniform(-0.5, 8.5, 200), rng.uniform(-0.5, 6.5, 200)], -1)
    for mode in ("bilinear", "nearest", "bicubic"):
        out = _np(tt.sample_image_2d(_t(img, device), _t(c, device), interpolation=mode))
        assert rel_l2(out, O.sample_image_2d(img, c, mode)) < 1e-12
    with pytest.raises(ValueError):
        tt.sample_image_2d(torch.zeros(4, device=device), torch.zeros(3, 2, device=device))
    with pytest.raises(ValueError):
        tt.sample_image_2d(torch.zeros(4, 4, device=device), torch.zeros(3, 2, device=device), "trilinear")
    with pytest.raises(RuntimeError):
        tt.sample_image_2d(torch.zeros(4, 4, device=device), torch.zeros(3, 2, device=device, dtype=torch.float64))


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_insert_into_image_2d_vs_golden(device, mode):
    G = golden("insert_into_image_2d")
    for key, v1, v2, shape, dt in (("real", "val_r1", "val_r2", (12, 10), torch.float32),
                                   ("complex", "val_c1", "val_c2", (12, 10), torch.complex64),
                                   ("multi", "val_m1", "val_m2", (2, 12, 10), torch.float32)):
        img = torch.zeros(shape, dtype=dt, device=device)
        img, w = tt.insert_into_image_2d(_t(G[v1], device), _t(G["coords1"], device), img, interpolation=mode)
        img2, w2 = tt.insert_into_image_2d(_t(G[v2], device), _t(G["coords2"], device), img, w,
                                           interpolation=mode)
        assert img2.data_ptr() == img.data_ptr() and w2.data_ptr() == w.data_ptr()  # in place
        assert rel_l2(_np(img2), G[f"{key}_{mode}_image"]) < INS_TOL
        assert rel_l2(_np(w2), G[f"{key}_{mode}_weights"]) < INS_TOL


def test_insert_into_image_2d_known_answers_and_types(device):
    # TII tests/test_image_interpolation_2d.py:65-80, :143-160, :163-190
    img, w = tt.insert_into_image_2d(torch.tensor([5.0], device=device), torch.tensor([[10.5, 14.5]], device=device),
                                     torch.zeros((28, 28), device=device))
    assert torch.allclose(img[10:12, 14:16], torch.full((2, 2), 1.25, device=device))
    img, w = tt.insert_into_image_2d(torch.tensor([5.0], device=device), torch.tensor([[10.7, 14.3]], device=device),
                                     torch.zeros((28, 28), device=device), interpolation="nearest")
    assert img[11, 14] == 5 and float(img.sum()) == 5.0
    for dtype in (torch.float32, torch.float64, torch.complex64, torch.complex128):
        image = torch.zeros((4, 4), dtype=dtype, device=device)
        weights = torch.zeros((4, 4), dtype=torch.float64, device=device)
        coords = torch.rand((3, 4, 2), device=device) * 3
        values = torch.ones((3, 4), dtype=dtype, device=device)
        for mode in ("bilinear", "nearest"):
            image, weights = tt.insert_into_image_2d(values, coords, image, weights, interpolation=mode)
            assert image.dtype == dtype and weights.dtype == torch.float64
        # float32 corner weights (as the reference computes them) sum to 1 within rounding
        assert abs(float(weights.sum()) - 24.0) < 1e-5
    with pytest.raises(ValueError):
        tt.insert_into_image_2d(torch.ones((3, 4), dtype=torch.complex128, device=device),
                                torch.rand((3, 4, 2), device=device),
                                torch.zeros((4, 4), dtype=torch.complex64, device=device))


# ---------------------------------------------------------------- 1-D samplers (on the 2-D kernels)


@pytest.mark.parametrize("mode", ["linear", "cubic", "nearest"])
def test_sample_image_1d_vs_oracle(device, mode):
    rng = np.random.default_rng(60)
    img = rng.standard_normal(40).astype(np.float32)
    mc = (rng.standard_normal((3, 40)) + 1j * rng.standard_normal((3, 40))).astype(np.complex64)
    c = rng.uniform(-2, 41, (5, 70)).astype(np.float32)
    c[0, :6] = [0.0, 39.0, 12.5, 3.5, 7.5, 20.0]  # edges and nearest ties
    out = tt.sample_image_1d(_t(img, device), _t(c, device), mode)
    assert out.shape == (5, 70)
    assert max_abs(_np(out), O.sample_image_1d(img, c, mode)) < 1e-5
    out = tt.sample_image_1d(_t(mc, device), _t(c, device), mode)
    assert out.shape == (5, 70, 3) and out.dtype == torch.complex64
    assert max_abs(_np(out), O.sample_image_1d(mc, c, mode)) < 1e-5


def test_sample_image_1d_reference_cases(device):
    """TII tests/test_image_interpolation_1d.py:9-63: integer coordinates of
    any (...) shape, complex and multichannel images."""
    rng = np.random.default_rng(61)
    coords = torch.tensor(rng.integers(0, 27, size=(6, 7, 8)), device=device)
    image = torch.rand(28, device=device)
    for mode in ("nearest", "linear", "cubic"):
        s = tt.sample_image_1d(image, coords, interpolation=mode)
        assert s.shape == (6, 7, 8)
        assert torch.allclose(s, image[coords])  # integer positions read the pixel
    cimg = torch.complex(torch.rand(28, device=device), torch.rand(28, device=device))
    assert tt.sample_image_1d(cimg, coords).shape == (6, 7, 8)
    assert tt.sample_image_1d(torch.rand(3, 28, device=device), coords).shape == (6, 7, 8, 3)
    assert tt.sample_image_1d(cimg.expand(3, 28).contiguous(), coords).shape == (6, 7, 8, 3)
    with pytest.raises(ValueError):
        tt.sample_image_1d(torch.rand(2, 3, 28, device=device), coords)


def test_insert_into_image_1d_reference_cases(device):
    """TII tests/test_image_interpolation_1d.py:66-195: known answers, in-place
    accumulation, multichannel shapes, dtype consistency and the dtype error."""
    image = torch.zeros(28, device=device)
    out, w = tt.insert_into_image_1d(torch.tensor([1.0], device=device), torch.tensor([14.28], device=device), image)
    want = torch.zeros(28, device=device)
    want[14], want[15] = 0.72, 0.28
    assert out is image and torch.allclose(out, want) and torch.allclose(w, want)
    image = torch.zeros(28, device=device)
    out, _ = tt.insert_into_image_1d(torch.tensor([5.0], device=device), torch.tensor([10.7], device=device), image,
                                     interpolation="nearest")
    want = torch.zeros(28, device=device)
    want[11] = 5
    assert torch.allclose(out, want)
    rng = np.random.default_rng(62)
    coords = torch.tensor(rng.integers(0, 27, size=(6, 7, 8)), device=device)
    for mode in ("linear", "nearest"):
        img3 = torch.zeros(3, 28, device=device)
        img3, w3 = tt.insert_into_image_1d(torch.ones(6, 7, 8, 3, device=device), coords, img3, interpolation=mode)
        assert img3.shape == (3, 28) and w3.shape == (28,) and img3[0, coords[0, 0, 0]] > 0
    for dtype in (torch.float32, torch.float64, torch.complex64, torch.complex128):
        img = torch.rand(4, dtype=dtype, device=device)
        wts = torch.zeros(4, dtype=torch.float64, device=device)
        cc = torch.tensor(rng.uniform(0, 3, size=(3, 4)), device=device)
        for mode in ("linear", "nearest"):
            img, wts = tt.insert_into_image_1d(torch.rand(3, 4, dtype=dtype, device=device), cc, img, wts, mode)
            assert img.dtype == dtype and wts.dtype == torch.float64
    with pytest.raises(ValueError):
        tt.insert_into_image_1d(torch.rand(3, 4, dtype=torch.complex128, device=device),
                                torch.zeros(3, 4, device=device), torch.rand(4, dtype=torch.complex64, device=device))


@pytest.mark.parametrize("mode", ["linear", "nearest"])
def test_insert_into_image_1d_vs_oracle(device, mode):
    rng = np.random.default_rng(63)
    vals = (rng.standard_normal((500, 2)) + 1j * rng.standard_normal((500, 2))).astype(np.complex64)
    c = rng.uniform(-1, 50, 500).astype(np.float32)
    c[:4] = [0.0, 49.0, 10.5, 11.5]
    img = torch.zeros(2, 50, dtype=torch.complex64, device=device)
    img, w = tt.insert_into_image_1d(_t(vals, device), _t(c, device), img, interpolation=mode)
    ri, rw = O.insert_into_image_1d(vals.astype(np.complex128), c, np.zeros((2, 50), np.complex128),
                                    interpolation=mode)
    assert max_abs(_np(img), ri) < 1e-4 and max_abs(_np(w), rw) < 1e-4
