"""Parity of the gfx950 central-slice kernels against the FP64 oracle and the
reference's golden vectors.  Tolerances (relative L2): extraction <= 1e-5,
insertion <= 1e-4 (float atomics reorder the sums), gradients <= 1e-4."""

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

import teamtomo_amd as tt
from conftest import golden, max_abs, rel_l2
from oracle import fourier_slice as O
from teamtomo_amd.synthetic import blob_volume, uniform_rotations, volume_rfft

pytestmark = pytest.mark.gpu

EXT_TOL = 1e-5
INS_TOL = 1e-4
GRAD_TOL = 1e-4


def _t(a, device, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(device)
    return t if dtype is None else t.to(dtype)


def _np(t):
    return t.detach().cpu().numpy()


# ----------------------------------------------------------------- extraction
def test_extract_64cubed_100rot_vs_oracle_and_golden(device):
    G = golden("extract_insert_64")
    V = _t(G["volume_rfft"], device)
    R = _t(G["rotations"], device)
    S = _np(tt.extract_central_slices_rfft_3d(V, (64, 64, 64), R))
    ref = O.extract_central_slices(G["volume_rfft"], (64, 64, 64), O.prepare_rotations(G["rotations"], False, 1e-8))
    print("extract 64^3 x 100 rel_l2 vs FP64:", rel_l2(S, ref), "max", max_abs(S, ref))
    assert rel_l2(S, ref) < EXT_TOL
    assert rel_l2(S, G["slices"]) < EXT_TOL


@pytest.mark.parametrize("name,shape,fmax", [
    ("odd_33", (33, 33, 33), None), ("fftfreq_max_03", (32, 32, 32), 0.3), ("image_shape_4x", (64, 64, 64), None),
])
def test_extract_golden_cases(device, name, shape, fmax):
    G = golden(name)
    S = _np(tt.extract_central_slices_rfft_3d(_t(G["volume_rfft"], device), shape, _t(G["rotations"], device),
                                              fftfreq_max=fmax))
    ref = O.extract_central_slices(G["volume_rfft"], shape, O.prepare_rotations(G["rotations"], False, 1e-8), fmax)
    assert rel_l2(S, ref) < EXT_TOL
    assert rel_l2(S, G["slices"]) < EXT_TOL
    assert np.array_equal(S == 0, ref == 0)


def test_extract_conjugate_pairs_and_tiebreak(device):
    G = golden("tiebreak_32")
    V = _t(G["volume_rfft"], device)
    SL = _np(tt.extract_central_slices_rfft_3d(V, (32, 32, 32), _t(G["left"], device)))
    SR = _np(tt.extract_central_slices_rfft_3d(V, (32, 32, 32), _t(G["right"], device)))
    scale = np.abs(G["volume_rfft"]).max()
    assert max_abs(SL, np.conj(SR)) < 1e-6 * scale
    assert rel_l2(SL, G["slices_left"]) < EXT_TOL
    # 100 random in-plane 180-degree pairs (reference test :189-232)
    rng = np.random.default_rng(0)
    ang = rng.uniform([0, 0, 0], [360, 180, 360], (100, 3))
    left = Rotation.from_euler("ZYZ", ang, degrees=True).as_matrix()
    right = left @ Rotation.from_euler("ZYZ", [0, 0, 180], degrees=True).as_matrix()
    SL = _np(tt.extract_central_slices_rfft_3d(V, (32, 32, 32), _t(left, device)))
    SR = _np(tt.extract_central_slices_rfft_3d(V, (32, 32, 32), _t(right, device)))
    assert max_abs(SL, np.conj(SR)) < 1e-6 * scale


def test_extract_x0_friedel_line(device):
    vol = blob_volume(64, 20, seed=3, device=device)
    V = volume_rfft(vol)
    S = _np(tt.extract_central_slices_rfft_3d(V, (64, 64, 64), uniform_rotations(10, 3, device=device)))
    line = S[:, :, 0]
    assert max_abs(line[:, 1:32], np.conj(line[:, -31:])[:, ::-1]) < 1e-6 * np.abs(line).max()


def test_extract_identity_exact(device):
    V = volume_rfft(torch.randn(32, 32, 32, generator=torch.Generator().manual_seed(2))).to(device)
    S = tt.extract_central_slices_rfft_3d(V, (32, 32, 32), torch.eye(3, device=device)[None])
    assert torch.equal(S[0], V[16])


def test_extract_shapes_dtypes_and_errors(device):
    V = volume_rfft(torch.randn(16, 16, 16)).to(device)
    R = torch.rand(4, 5, 3, 3, device=device)
    assert tt.extract_central_slices_rfft_3d(V, (16, 16, 16), R).shape == (4, 5, 16, 9)
    R64 = uniform_rotations(3, 1, device=device, dtype=torch.float64)
    assert tt.extract_central_slices_rfft_3d(V, (16, 16, 16), R64).dtype == torch.complex64
    empty = tt.extract_central_slices_rfft_3d(V, (16, 16, 16), torch.zeros(0, 3, 3, device=device))
    assert empty.shape == (0, 16, 9)
    Rf = uniform_rotations(3, 2, device=device)
    Rf[0, 0, 1] = 1e-9
    before = Rf.clone()
    tt.extract_central_slices_rfft_3d(V, (16, 16, 16), Rf)
    assert torch.equal(Rf, before)  # no in-place clamp of the caller's tensor
    with pytest.raises(AssertionError):
        tt.extract_central_slices_rfft_3d(V[None], (16, 16, 16), Rf)
    with pytest.raises(RuntimeError):
        tt.extract_central_slices_rfft_3d(V.cpu(), (16, 16, 16), Rf.cpu())


def test_empty_batches_give_zero_outputs(device):
    """No slices: insertion returns a zero volume and zero weights, and the
    volume gradient of an empty extraction is zero (the outputs are freshly
    allocated, so nothing may be left unwritten)."""
    n = 16
    for dtype in (torch.complex64, torch.complex128):
        for _ in range(2):  # fresh allocations may reuse the same (dirty) memory
            junk = torch.full((4 * n * n * n,), 7.0, device=device)
            del junk
            vol, w = tt.insert_central_slices_rfft_3d(torch.zeros(0, n, n // 2 + 1, dtype=dtype, device=device),
                                                      (n, n, n), torch.zeros(0, 3, 3, device=device))
            assert vol.shape == (n, n, n // 2 + 1) and vol.dtype == dtype
            assert not bool(vol.abs().amax()) and not bool(w.abs().amax())
    V = volume_rfft(torch.randn(n, n, n)).to(device).requires_grad_(True)
    S = tt.extract_central_slices_rfft_3d(V, (n, n, n), torch.zeros(0, 3, 3, device=device))
    torch.real(torch.sum(S)).backward()
    assert V.grad is not None and not bool(V.grad.abs().amax())
    img = torch.zeros(8, 8, device=device)
    out, wts = tt.insert_into_image_2d(torch.zeros(0, device=device), torch.zeros(0, 2, device=device), img)
    assert out is img and not bool(wts.abs().amax())
    assert tt.sample_image_3d(torch.randn(4, 4, 4, device=device), torch.zeros(0, 3, device=device)).shape == (0,)


def test_prepare_rotations_kernel_matches_torch_steps(device):
    """tt_prepare_rotations_*: float32 cast, |R| < tol -> 0 (in float32) and
    the xyz -> zyx flip in one pass, equal to the reference's torch steps
    (TFS extract :147-154) for float32 and float64 input, any leading dims."""
    from teamtomo_amd.fourier_slice import _prepare_rotations

    g = torch.Generator().manual_seed(31)
    R = torch.randn(4, 5, 3, 3, generator=g, dtype=torch.float64)
    R[0, 0, 0, 1] = 5e-9  # below the tolerance
    R[0, 1, 2, 2] = -2e-8  # above it
    R[1, 0, 1, 0] = 1e-8 * (1 - 1e-9)  # rounds to float32(1e-8): not below it in float32
    for dt in (torch.float32, torch.float64):
        for zyx in (False, True):
            for tol in (None, 1e-8):
                Rd = R.to(dt).to(device)
                got = _prepare_rotations(Rd, device, zyx, tol)
                want = R.to(dt).to(torch.float32)
                if tol is not None:
                    want = torch.where(torch.abs(want) < tol, torch.zeros(()), want)
                if not zyx:
                    want = torch.flip(want, dims=(-2, -1))
                assert got.dtype == torch.float32 and got.shape == R.shape
                assert torch.equal(got.cpu(), want), (dt, zyx, tol)
    # a rotation that requires grad takes the differentiable torch steps
    Rg = R.to(torch.float32).to(device).requires_grad_(True)
    _prepare_rotations(Rg, device, False, 1e-8).sum().backward()
    assert Rg.grad is not None


def test_projections_of_empty_batches(device):
    """No poses / no images: empty projections and a zero backprojection
    (hipFFT refuses empty batches, so these skip the FFT)."""
    vol = torch.rand(16, 16, 16, device=device)
    none = torch.zeros(0, 3, 3, device=device)
    assert tt.project_3d_to_2d(vol, none).shape == (0, 16, 16)
    rec = tt.backproject_2d_to_3d(torch.zeros(0, 16, 16, device=device), none)
    assert rec.shape == (16, 16, 16) and not bool(rec.abs().amax())
    assert tt.project_2d_to_1d(torch.rand(16, 16, device=device), torch.zeros(0, 2, 2, device=device)).shape == (0, 16)
    # the empty result has the shape a non-empty batch would have, odd padded boxes included
    # (irfftn without s gives 2 (Wr - 1) = N - 1 for an odd N)
    one3, one2 = torch.eye(3, device=device)[None], torch.eye(2, device=device)[None]
    for d in (16, 33):
        for pad in (1.0, 2.0):
            v = torch.rand(d, d, d, device=device)
            full = tt.project_3d_to_2d(v, one3, pad_factor=pad).shape[1:]
            assert tt.project_3d_to_2d(v, none, pad_factor=pad).shape == (0,) + full, (d, pad, full)
            mc = torch.rand(2, d, d, d, device=device)
            assert tt.project_3d_to_2d_multichannel(mc, none, pad_factor=pad).shape == (0, 2) + full
            im = torch.rand(d, d, device=device)
            full1 = tt.project_2d_to_1d(im, one2, pad_factor=pad).shape[1:]
            assert tt.project_2d_to_1d(im, torch.zeros(0, 2, 2, device=device), pad_factor=pad).shape == (0,) + full1


# ------------------------------------------------------------------ insertion
def test_insert_64cubed_100rot_vs_oracle_and_golden(device):
    G = golden("extract_insert_64")
    Vi, Wi = tt.insert_central_slices_rfft_3d(_t(G["slices"], device), (64, 64, 64), _t(G["rotations"], device))
    rV, rW = O.insert_central_slices(G["slices"], (64, 64, 64), O.prepare_rotations(G["rotations"], False, None))
    print("insert 64^3 x 100 rel_l2 vs FP64:", rel_l2(_np(Vi), rV), rel_l2(_np(Wi), rW))
    assert Vi.dtype == torch.complex64 and Wi.dtype == torch.float32
    assert rel_l2(_np(Vi), rV) < INS_TOL and rel_l2(_np(Wi), rW) < INS_TOL
    assert rel_l2(_np(Vi), G["ins_volume"]) < INS_TOL and rel_l2(_np(Wi), G["ins_weights"]) < INS_TOL


@pytest.mark.parametrize("name,shape,fmax", [("odd_33", (33, 33, 33), None), ("fftfreq_max_03", (32, 32, 32), 0.3)])
def test_insert_golden_cases(device, name, shape, fmax):
    G = golden(name)
    Vi, Wi = tt.insert_central_slices_rfft_3d(_t(G["slices"], device), shape, _t(G["rotations"], device),
                                              fftfreq_max=fmax)
    rV, rW = O.insert_central_slices(G["slices"], shape, O.prepare_rotations(G["rotations"], False, None), fmax)
    assert rel_l2(_np(Vi), rV) < INS_TOL and rel_l2(_np(Wi), rW) < INS_TOL
    assert rel_l2(_np(Vi), G["ins_volume"]) < INS_TOL


def test_insert_friedel_rotations_golden(device):
    G = golden("tiebreak_32")
    R = _t(G["friedel_rotations"], device)
    Vi, Wi = tt.insert_central_slices_rfft_3d(_t(G["friedel_slices"], device), (32, 32, 32), R)
    assert rel_l2(_np(Vi), G["friedel_volume"]) < INS_TOL
    assert max_abs(_np(Wi), G["friedel_weights"]) < 1e-5
    x0 = _np(Vi)[:, :, 0]
    zz, yy = np.meshgrid(np.arange(32), np.arange(32), indexing="ij")
    err = np.abs(x0 - np.conj(x0[(32 - zz) % 32, (32 - yy) % 32]))
    err[16, 16] = 0
    assert err.mean() < 1e-5


def test_insert_complex128(device):
    rng = np.random.default_rng(3)
    d = rng.standard_normal((5, 24, 13)) + 1j * rng.standard_normal((5, 24, 13))
    R = Rotation.random(5, random_state=3).as_matrix()
    Vi, Wi = tt.insert_central_slices_rfft_3d(_t(d, device), (24, 24, 24), _t(R, device))
    assert Vi.dtype == torch.complex128 and Wi.dtype == torch.float64
    rV, rW = O.insert_central_slices(d, (24, 24, 24), O.prepare_rotations(R, False, None))
    assert rel_l2(_np(Vi), rV) < 1e-6 and rel_l2(_np(Wi), rW) < 1e-6


def test_insert_identity_known_answer(device):
    N = 16
    d = torch.randn(1, N, N // 2 + 1, dtype=torch.complex64, generator=torch.Generator().manual_seed(4))
    Vi, Wi = tt.insert_central_slices_rfft_3d(d.to(device), (N, N, N), torch.eye(3, device=device)[None])
    rV, rW = O.insert_central_slices(d.numpy(), (N, N, N), O.prepare_rotations(np.eye(3)[None], False, None))
    assert max_abs(_np(Vi), rV) < 1e-6 and max_abs(_np(Wi), rW) == 0.0


# ------------------------------------------------------------------ autograd
def test_extract_backward_vs_oracle_and_golden(device):
    G = golden("autograd_8")
    V = _t(G["volume_rfft"], device).requires_grad_(True)
    R = _t(G["ext_rotations"], device).requires_grad_(True)
    S = tt.extract_central_slices_rfft_3d(V, (8, 8, 8), R)
    loss = torch.real(torch.sum(torch.conj(_t(G["ext_target"], device)) * S))
    loss.backward()
    assert rel_l2(_np(V.grad), G["ext_grad_volume"]) < GRAD_TOL
    assert rel_l2(_np(R.grad), G["ext_grad_rotations"]) < GRAD_TOL
    assert R.grad.dtype == torch.float64


def test_insert_backward_vs_golden(device):
    G = golden("autograd_8")
    d = _t(G["ins_slices"], device).requires_grad_(True)
    R = _t(G["ins_rotations"], device).requires_grad_(True)
    Vi, Wi = tt.insert_central_slices_rfft_3d(d, (8, 8, 8), R)
    loss = torch.real(torch.sum(torch.conj(_t(G["ins_target_volume"], device)) * Vi)) + torch.sum(
        _t(G["ins_target_weights"], device) * Wi)
    loss.backward()
    assert rel_l2(_np(d.grad), G["ins_grad_slices"]) < GRAD_TOL
    assert rel_l2(_np(R.grad), G["ins_grad_rotations"]) < GRAD_TOL


def test_extract_adjoint_vs_oracle_64cubed_100rot(device):
    G = golden("extract_insert_64")
    rng = np.random.default_rng(5)
    T = (rng.standard_normal(G["slices"].shape) + 1j * rng.standard_normal(G["slices"].shape)).astype(np.complex64)
    V = _t(G["volume_rfft"], device).requires_grad_(True)
    S = tt.extract_central_slices_rfft_3d(V, (64, 64, 64), _t(G["rotations"], device))
    S.backward(_t(T, device))
    Rz = O.prepare_rotations(G["rotations"], False, 1e-8)
    ref = O.extract_adjoint(T, (64, 64, 33), (64, 64, 64), Rz)
    print("adjoint 64^3 x 100 rel_l2 vs FP64:", rel_l2(_np(V.grad), ref))
    assert rel_l2(_np(V.grad), ref) < INS_TOL


def test_rotation_grad_vs_finite_differences(device):
    """Kernel rotation gradient vs FP64 central differences of the oracle."""
    rng = np.random.default_rng(6)
    vol = O.volume_rfft(rng.standard_normal((16, 16, 16)))
    R = Rotation.random(4, random_state=6).as_matrix()
    T = rng.standard_normal((4, 16, 9)) + 1j * rng.standard_normal((4, 16, 9))
    Rt = _t(R, device).requires_grad_(True)
    S = tt.extract_central_slices_rfft_3d(_t(vol.astype(np.complex64), device), (16, 16, 16), Rt)
    torch.real(torch.sum(torch.conj(_t(T.astype(np.complex64), device)) * S)).backward()
    Rz = O.prepare_rotations(R, False, 1e-8).astype(np.float64)
    volc = vol.astype(np.complex64).astype(np.complex128)
    loss = lambda Rm: np.real(np.vdot(T.astype(np.complex64), O.extract_central_slices(volc, (16, 16, 16), Rm)))
    h = 1e-7
    fd = np.zeros_like(Rz)
    for idx in np.ndindex(*Rz.shape):
        Rp, Rm = Rz.copy(), Rz.copy()
        Rp[idx] += h
        Rm[idx] -= h
        fd[idx] = (loss(Rp) - loss(Rm)) / (2 * h)
    fd_xyz = O.rotation_grad_to_xyz(fd, R, False, 1e-8)
    print("rotation grad rel_l2 vs FD:", rel_l2(_np(Rt.grad), fd_xyz))
    assert rel_l2(_np(Rt.grad), fd_xyz) < GRAD_TOL


# ------------------------------------------------------------ generic interp
@pytest.mark.parametrize("mode", ["trilinear", "nearest"])
def test_sample_image_3d_golden(device, mode):
    G = golden("sample_image_3d")
    for key in ("real", "complex", "multi"):
        out = _np(tt.sample_image_3d(_t(G["image_" + key], device), _t(G["coordinates"], device), mode))
        assert out.shape == G[f"{key}_{mode}"].shape
        assert rel_l2(out, G[f"{key}_{mode}"]) < EXT_TOL
        if mode == "nearest":
            assert max_abs(out, G[f"{key}_{mode}"]) == 0.0


@pytest.mark.parametrize("mode", ["trilinear", "nearest"])
def test_insert_into_image_3d_golden(device, mode):
    G = golden("insert_into_image_3d")
    for key, kk, shape, dt in (("r", "real", (12, 10, 9), torch.float32), ("c", "complex", (12, 10, 9), torch.complex64),
                               ("m", "multi", (2, 12, 10, 9), torch.float32)):
        img = torch.zeros(shape, dtype=dt, device=device)
        img, w = tt.insert_into_image_3d(_t(G[f"val_{key}1"], device), _t(G["coords1"], device), img, interpolation=mode)
        img2, w2 = tt.insert_into_image_3d(_t(G[f"val_{key}2"], device), _t(G["coords2"], device), img, w,
                                           interpolation=mode)
        assert img2 is img and w2 is w  # in place
        assert rel_l2(_np(img), G[f"{kk}_{mode}_image"]) < INS_TOL
        assert rel_l2(_np(w), G[f"{kk}_{mode}_weights"]) < INS_TOL


def test_insert_into_image_known_answers(device):
    img = torch.zeros(28, 28, 28, device=device)
    img, w = tt.insert_into_image_3d(torch.tensor([5.0], device=device),
                                     torch.tensor([[10.5, 14.5, 18.5]], device=device), img)
    assert torch.allclose(img[10:12, 14:16, 18:20], torch.full((2, 2, 2), 5 / 8, device=device))
    img = torch.zeros(28, 28, 28, device=device)
    img, w = tt.insert_into_image_3d(torch.tensor([5.0], device=device),
                                     torch.tensor([[10.7, 14.3, 18.9]], device=device), img, interpolation="nearest")
    assert img[11, 14, 19] == 5 and img.sum() == 5
    with pytest.raises(ValueError):
        tt.insert_into_image_3d(torch.zeros(3, 4, 5, dtype=torch.complex128, device=device),
                                torch.zeros(3, 4, 5, 3, device=device),
                                torch.zeros(4, 4, 4, dtype=torch.complex64, device=device))
    w64 = torch.zeros(4, 4, 4, dtype=torch.float64, device=device)
    img = torch.zeros(4, 4, 4, dtype=torch.float64, device=device)
    _, w64 = tt.insert_into_image_3d(torch.ones(3, dtype=torch.float64, device=device),
                                     torch.full((3, 3), 1.25, device=device), img, w64)
    assert w64.dtype == torch.float64 and abs(float(w64.sum()) - 3.0) < 1e-12


# ------------------------------------------------------------- project / backproject
def test_project_backproject_golden(device):
    G = golden("project_backproject_24")
    P = _np(tt.project_3d_to_2d(_t(G["volume"], device), _t(G["rotations"], device)))
    assert rel_l2(P, G["projections"]) < EXT_TOL
    B = _np(tt.backproject_2d_to_3d(_t(G["projections"], device), _t(G["rotations"], device)))
    assert rel_l2(B, G["backprojection"]) < INS_TOL


def test_project_rotation_centre(device):
    vol = torch.zeros(32, 32, 32, device=device)
    vol[16, 16, 16] = 1
    P = tt.project_3d_to_2d(vol, uniform_rotations(100, 7, device=device))
    idx = P.reshape(100, -1).argmax(dim=-1)
    assert torch.all(idx == 16 * 32 + 16)


def _fsc(a, b):
    fa, fb = np.fft.fftshift(np.fft.fftn(a)), np.fft.fftshift(np.fft.fftn(b))
    n = a.shape[0]
    ax = np.arange(n) - n // 2
    r = np.sqrt(sum(g ** 2 for g in np.meshgrid(ax, ax, ax, indexing="ij")))
    shells = np.rint(r).astype(int)
    out = []
    for s in range(n // 2 + 1):
        m = shells == s
        num = np.real(np.sum(fa[m] * np.conj(fb[m])))
        den = np.sqrt(np.sum(np.abs(fa[m]) ** 2) * np.sum(np.abs(fb[m]) ** 2))
        out.append(num / den if den > 0 else 1.0)
    return np.array(out)


def test_projection_backprojection_cycle(device):
    cube = torch.zeros(32, 32, 32)
    cube[8:24, 8:24, 8:24] = 1
    cube[16, 16, 16] = 32
    R = uniform_rotations(1500, 8, device=device)
    P = tt.project_3d_to_2d(cube.to(device), R)
    rec = tt.backproject_2d_to_3d(P, R)
    f = _fsc(cube.numpy(), _np(rec))
    assert np.all(f[-10:] > 0.99)


def test_backproject_dtypes_and_leading_dims(device):
    imgs = torch.rand(10, 28, 28, dtype=torch.float64, device=device)
    R = uniform_rotations(10, 9, device=device, dtype=torch.float64)
    assert tt.backproject_2d_to_3d(imgs, R).dtype == torch.float64
    cube = torch.zeros(32, 32, 32, device=device)
    R = torch.rand(4, 5, 3, 3, device=device)
    P = tt.project_3d_to_2d(cube, R)
    assert P.shape == (4, 5, 32, 32)
    assert tt.backproject_2d_to_3d(P, R).shape == (32, 32, 32)


def test_multichannel_shapes(device):
    vols = torch.rand(4, 10, 10, 10, device=device)
    R = uniform_rotations(8, 10, device=device)
    P = tt.project_3d_to_2d_multichannel(vols, R)
    assert P.shape == (8, 4, 10, 10)
    assert tt.backproject_2d_to_3d_multichannel(P, R).shape == (4, 10, 10, 10)


# ------------------------------------------------------ scatter variants agree
@pytest.fixture
def scatter_path():
    yield tt.set_scatter_path
    tt.set_scatter_path("auto")


def _scatter_variants(path_setter):
    """(name, enter) pairs: the brick engine, the binned scatter and the
    pixel-driven atomic scatter."""
    def use(kind):
        def enter():
            path_setter("binned" if kind == "binned" else "brick")
            tt.set_pixel_atomic(kind == "atomic")
        return enter
    return [(k, use(k)) for k in ("brick", "binned", "atomic")]


@pytest.mark.parametrize("shape,image_shape,fmax", [
    ((48, 48, 48), (48, 48, 48), None), ((33, 33, 33), (33, 33, 33), 0.3), ((20, 20, 20), (64, 64, 64), None),
])
def test_scatter_variants_match_oracle(device, scatter_path, shape, image_shape, fmax):
    """Brick-engine LDS scatter, binned scatter and pixel-driven atomic scatter
    against the FP64 oracle and each other, for the adjoint and the insertion,
    incl. odd sizes and image_shape != volume."""
    n = 40
    R = uniform_rotations(n, 11, device=device)
    Vr = volume_rfft(torch.randn(*shape, generator=torch.Generator().manual_seed(12))).to(device)
    Hi, Wri = image_shape[1], image_shape[2] // 2 + 1
    g = torch.randn(n, Hi, Wri, dtype=torch.complex64, generator=torch.Generator().manual_seed(13)).to(device)
    d = torch.randn(n, shape[1], shape[2] // 2 + 1, dtype=torch.complex64,
                    generator=torch.Generator().manual_seed(14)).to(device)
    ref = O.extract_adjoint(_np(g), tuple(Vr.shape), image_shape, O.prepare_rotations(_np(R), False, 1e-8), fmax)
    rV, rW = O.insert_central_slices(_np(d), shape, O.prepare_rotations(_np(R), False, None), fmax)
    outs, ins = {}, {}
    for name, enter in _scatter_variants(scatter_path):
        enter()
        try:
            V = Vr.clone().requires_grad_(True)
            tt.extract_central_slices_rfft_3d(V, image_shape, R, fftfreq_max=fmax).backward(g)
            outs[name] = _np(V.grad)
            if image_shape == shape:
                ins[name] = [_np(x) for x in tt.insert_central_slices_rfft_3d(d, shape, R, fftfreq_max=fmax)]
        finally:
            tt.set_pixel_atomic(False)
        e_a = rel_l2(outs[name], ref)
        e_i = (rel_l2(ins[name][0], rV), rel_l2(ins[name][1], rW)) if name in ins else None
        print(f"{name}: adjoint rel_l2 {e_a:.2e}, insertion (volume, weights) {e_i}")
        assert e_a < INS_TOL
        if e_i is not None:
            assert e_i[0] < INS_TOL and e_i[1] < INS_TOL
    # the two fixed-point scatters: FP32 coordinate roundings apart, nothing more
    assert rel_l2(outs["brick"], outs["binned"]) < 1e-5
    if ins:
        assert rel_l2(ins["brick"][0], ins["binned"][0]) < 1e-5 and rel_l2(ins["brick"][1], ins["binned"][1]) < 1e-5


def test_split_bricks_adjoint_and_insert_vs_oracle(device):
    """1024 rotations at 64^3: the planner splits the bricks near DC into
    rotation segments (partial tiles combined by split_combine_kernel); the
    adjoint and the insertion still match the FP64 C oracle and the
    pixel-driven atomic scatter."""
    from oracle import c_oracle as C

    n, k = 64, 1024
    R = uniform_rotations(k, 21, device=device)
    Vr = volume_rfft(blob_volume(n, 30, seed=22, device=device))
    g = torch.complex(torch.randn(k, n, n // 2 + 1, generator=torch.Generator().manual_seed(23)),
                      torch.randn(k, n, n // 2 + 1, generator=torch.Generator().manual_seed(24))).to(device)
    Rz = O.prepare_rotations(_np(R), False, 1e-8)
    ref = C.extract_adjoint(_np(g), tuple(Vr.shape), (n, n, n), Rz)
    rV, rW = C.insert(_np(g), (n, n, n), O.prepare_rotations(_np(R), False, None))
    for atomic in (False, True):
        tt.set_pixel_atomic(atomic)
        try:
            V = Vr.clone().requires_grad_(True)
            tt.extract_central_slices_rfft_3d(V, (n, n, n), R).backward(g)
            Vi, Wi = tt.insert_central_slices_rfft_3d(g, (n, n, n), R)
        finally:
            tt.set_pixel_atomic(False)
        assert rel_l2(_np(V.grad), ref) < INS_TOL, atomic
        assert rel_l2(_np(Vi), rV) < INS_TOL and rel_l2(_np(Wi), rW) < INS_TOL, atomic


@pytest.mark.parametrize("path", ["brick", "binned"])
def test_deterministic_scatter_is_bitwise_reproducible(device, scatter_path, path):
    scatter_path(path)
    n = 64
    R = uniform_rotations(n, 15, device=device)
    V0 = volume_rfft(torch.randn(40, 40, 40, generator=torch.Generator().manual_seed(16))).to(device)
    g = torch.randn(n, 40, 21, dtype=torch.complex64, generator=torch.Generator().manual_seed(17)).to(device)
    tt.set_deterministic(True)
    try:
        grads, ins = [], []
        for _ in range(3):
            V = V0.clone().requires_grad_(True)
            tt.extract_central_slices_rfft_3d(V, (40, 40, 40), R).backward(g)
            grads.append(V.grad.clone())
            ins.append(tt.insert_central_slices_rfft_3d(g, (40, 40, 40), R))
    finally:
        tt.set_deterministic(False)
    assert all(torch.equal(grads[0], x) for x in grads[1:])
    assert all(torch.equal(ins[0][0], x[0]) and torch.equal(ins[0][1], x[1]) for x in ins[1:])
    ref = O.extract_adjoint(_np(g), (40, 40, 21), (40, 40, 40), O.prepare_rotations(_np(R), False, 1e-8))
    assert rel_l2(_np(grads[0]), ref) < 1e-5
    rV, rW = O.insert_central_slices(_np(g), (40, 40, 40), O.prepare_rotations(_np(R), False, None))
    assert rel_l2(_np(ins[0][0]), rV) < 1e-5 and rel_l2(_np(ins[0][1]), rW) < 1e-5


@pytest.mark.parametrize("path", ["brick", "binned"])
def test_scatter_accumulate_flag(device, scatter_path, path):
    """Without TT_FLAG_OVERWRITE the C ABI adds into the caller's buffers."""
    from teamtomo_amd import _lib

    scatter_path(path)
    n = 8
    R = uniform_rotations(n, 18, device=device)
    Rz = torch.flip(R, dims=(-2, -1)).contiguous()
    d = torch.randn(n, 24, 13, dtype=torch.complex64, generator=torch.Generator().manual_seed(19)).to(device)
    base_v = torch.randn(24, 24, 13, dtype=torch.complex64, generator=torch.Generator().manual_seed(20)).to(device)
    base_w = torch.rand(24, 24, 13, generator=torch.Generator().manual_seed(21)).to(device)
    vol, w = base_v.clone(), base_w.clone()
    _lib.call("tt_insert_c64", _lib.ptr(d), 24, 24, 24, _lib.ptr(Rz), n, -1.0, _lib.ptr(vol), _lib.ptr(w), 0,
              _lib.stream_of(d))
    Vi, Wi = tt.insert_central_slices_rfft_3d(d, (24, 24, 24), R)
    assert rel_l2(_np(vol), _np(base_v + Vi)) < 1e-6 and rel_l2(_np(w), _np(base_w + Wi)) < 1e-6
    # the adjoint adds into grad_volume the same way
    gv = base_v.clone()
    Rz8 = tt.fourier_slice._prepare_rotations(R, device, False, 1e-8).reshape(-1, 3, 3).contiguous()
    _lib.call("tt_extract_bwd_c64", _lib.ptr(d), _lib.ptr(base_v), 24, 24, 13, 24, 24, 24, _lib.ptr(Rz8), n,
              -1.0, 1e-8, _lib.ptr(gv), None, 0, _lib.stream_of(d))
    V = base_v.clone().requires_grad_(True)
    tt.extract_central_slices_rfft_3d(V, (24, 24, 24), R).backward(d)
    assert rel_l2(_np(gv), _np(base_v + V.grad)) < 1e-6


# ------------------------------------------------------ brick gather (LDS)
@pytest.fixture
def gather_path():
    yield tt.set_gather_path
    tt.set_gather_path("auto")


@pytest.mark.parametrize("path", ["pixel", "brick"])
@pytest.mark.parametrize("name,shape,fmax", [
    ("extract_insert_64", (64, 64, 64), None), ("odd_33", (33, 33, 33), None),
    ("fftfreq_max_03", (32, 32, 32), 0.3), ("image_shape_4x", (64, 64, 64), None),
])
def test_extract_paths_vs_oracle_and_golden(device, gather_path, path, name, shape, fmax):
    gather_path(path)
    G = golden(name)
    S = _np(tt.extract_central_slices_rfft_3d(_t(G["volume_rfft"], device), shape, _t(G["rotations"], device),
                                              fftfreq_max=fmax))
    ref = O.extract_central_slices(G["volume_rfft"], shape, O.prepare_rotations(G["rotations"], False, 1e-8), fmax)
    print(f"{path} {name}: rel_l2 vs FP64 {rel_l2(S, ref):.2e}, vs golden {rel_l2(S, G['slices']):.2e}")
    assert rel_l2(S, ref) < EXT_TOL
    assert rel_l2(S, G["slices"]) < EXT_TOL
    assert np.array_equal(S == 0, ref == 0)


def _exact_rotations():
    m = [np.eye(3), [[0, -1, 0], [1, 0, 0], [0, 0, 1]], [[-1, 0, 0], [0, -1, 0], [0, 0, 1]],
         [[1, 0, 0], [0, 0, -1], [0, 1, 0]], [[0, 0, 1], [0, 1, 0], [-1, 0, 0]]]
    return np.asarray(m, dtype=np.float32)


@pytest.mark.parametrize("n,image,nrot,fmax", [
    (32, 32, 300, None), (33, 33, 200, None), (32, 64, 100, None), (48, 48, 300, 0.4), (256, 256, 96, None),
])
def test_brick_gather_equals_pixel_gather(device, gather_path, n, image, nrot, fmax):
    """Both gathers evaluate the same FP64 coordinates, cells and weights, so
    the samples agree to the last bit; exact-geometry rotations (samples on
    voxel faces and brick boundaries) are included."""
    V = volume_rfft(blob_volume(n, 30, seed=n, device=device))
    R = torch.cat([_t(_exact_rotations(), device), uniform_rotations(nrot, n, device=device)])
    gather_path("pixel")
    a = _np(tt.extract_central_slices_rfft_3d(V, (image,) * 3, R, fftfreq_max=fmax))
    gather_path("brick")
    b = _np(tt.extract_central_slices_rfft_3d(V, (image,) * 3, R, fftfreq_max=fmax))
    diff = max_abs(a, b)
    print(f"n={n} image={image} nrot={nrot}: max |brick - pixel| = {diff:.3e}, bitwise {np.array_equal(a, b)}")
    assert diff <= 1e-6 * np.abs(a).max()
    assert np.array_equal(a == 0, b == 0)


def test_brick_gather_full_size_matches_pixel_and_oracle_subset(device, gather_path):
    """256^3 x 4096 rotations (the benchmark workload): brick and pixel gathers
    agree everywhere; 8 slices are checked against the C FP64 oracle."""
    from oracle import c_oracle as C

    V = volume_rfft(blob_volume(256, 200, seed=2, device=device))
    R = uniform_rotations(4096, 0, device=device)
    gather_path("brick")
    b = tt.extract_central_slices_rfft_3d(V, (256, 256, 256), R)
    gather_path("pixel")
    a = tt.extract_central_slices_rfft_3d(V, (256, 256, 256), R)
    diff = float((a - b).abs().max())
    scale = float(a.abs().max())
    print(f"full size: max |brick - pixel| = {diff:.3e} (scale {scale:.3e})")
    assert diff <= 1e-6 * scale
    idx = [0, 1, 517, 1024, 2047, 3000, 4094, 4095]
    Rn = O.prepare_rotations(R[idx].cpu().numpy(), False, 1e-8)
    ref = C.extract(V.cpu().numpy(), (256, 256, 256), Rn)
    sub = _np(b[idx])
    print("full size subset rel_l2 vs FP64:", rel_l2(sub, ref))
    assert rel_l2(sub, ref) < EXT_TOL


@pytest.mark.parametrize("n,img,fmax,nrot,seed", [
    (33, 33, None, 300, 81), (64, 64, 0.3, 257, 82), (32, 64, None, 200, 83), (48, 48, None, 1024, 84),
])
def test_extraction_writes_every_output_pixel(device, n, img, fmax, nrot, seed):
    """The zero pass and the brick gather together write every pixel of the
    output: buffers filled with NaN beforehand (the C ABI writes into the
    caller's buffer) hold no NaN afterwards, for the one-channel and the
    two-channel gather, with odd and padded image shapes and a disc mask;
    the values match the FP64 oracle."""
    from teamtomo_amd import _lib
    from teamtomo_amd.fourier_slice import _prepare_rotations

    V = volume_rfft(blob_volume(n, 30, seed=seed, device=device))
    V2 = torch.stack([V, volume_rfft(blob_volume(n, 30, seed=seed + 100, device=device))]).contiguous()
    R = uniform_rotations(nrot, seed, device=device)
    Rz = _prepare_rotations(R, device, False, 1e-8).reshape(-1, 3, 3).contiguous()
    D, H, Wr = V.shape
    wr = img // 2 + 1
    fm = -1.0 if fmax is None else float(fmax)
    st = _lib.stream_of(V)
    out = torch.full((nrot, img, wr), float("nan"), dtype=torch.complex64, device=device)
    _lib.call("tt_extract_c64", _lib.ptr(V), D, H, Wr, img, img, img, _lib.ptr(Rz), nrot, fm, 1e-8, _lib.ptr(out), st)
    out2 = torch.full((nrot, 2, img, wr), float("nan"), dtype=torch.complex64, device=device)
    _lib.call("tt_extract_mc_c64", _lib.ptr(V2), 2, D, H, Wr, img, img, img, _lib.ptr(Rz), nrot, fm, 1e-8,
              _lib.ptr(out2), st)
    torch.cuda.synchronize()
    assert not bool(torch.isnan(torch.view_as_real(out)).any()), "unwritten pixels (one channel)"
    assert not bool(torch.isnan(torch.view_as_real(out2)).any()), "unwritten pixels (two channels)"
    assert torch.equal(out2[:, 0], out)
    ref = O.extract_central_slices(_np(V), (img, img, img), O.prepare_rotations(_np(R), False, 1e-8), fmax)
    e = rel_l2(_np(out), ref)
    print(f"{n}^3 -> image {img}, fmax {fmax}, {nrot} rotations: rel_l2 {e:.2e}")
    assert e < EXT_TOL
    assert np.array_equal(_np(out) == 0, ref == 0)
