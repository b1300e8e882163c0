"""Pins of the FP64 oracle (CPU only).

Each oracle function is checked against (a) the golden vectors written from
the reference implementation (tests/golden/make_golden.py) and (b) closed
forms / invariants chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails.  The reference computes in float32 and the oracle in
float64, so golden comparisons allow 1e-5 relative L2 (observed <= 3e-6).
"""

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from conftest import golden, max_abs, rel_l2
from oracle import fourier_slice as O

GOLD_TOL = 1e-5


def _rng_volume_rfft(n, seed):
    rng = np.random.default_rng(seed)
    return O.volume_rfft(rng.standard_normal((n, n, n)))


def _rotations(n, seed):
    return Rotation.random(n, random_state=seed).as_matrix()


# ----------------------------------------------------------------- golden vectors
def test_golden_extract_insert_64():
    G = golden("extract_insert_64")
    S = O.extract_central_slices(G["volume_rfft"], (64, 64, 64), O.prepare_rotations(G["rotations"], False, 1e-8))
    assert rel_l2(S, G["slices"]) < GOLD_TOL
    V, W = O.insert_central_slices(G["slices"], (64, 64, 64), O.prepare_rotations(G["rotations"], False, None))
    assert rel_l2(V, G["ins_volume"]) < GOLD_TOL
    assert rel_l2(W, G["ins_weights"]) < GOLD_TOL


@pytest.mark.parametrize("name,shape,fmax", [("odd_33", (33, 33, 33), None), ("fftfreq_max_03", (32, 32, 32), 0.3)])
def test_golden_odd_and_fftfreq_max(name, shape, fmax):
    G = golden(name)
    S = O.extract_central_slices(G["volume_rfft"], shape, O.prepare_rotations(G["rotations"], False, 1e-8), fmax)
    assert rel_l2(S, G["slices"]) < GOLD_TOL
    # masked pixels are exactly zero in both
    assert np.array_equal(S == 0, G["slices"] == 0)
    V, W = O.insert_central_slices(G["slices"], shape, O.prepare_rotations(G["rotations"], False, None), fmax)
    assert rel_l2(V, G["ins_volume"]) < GOLD_TOL
    assert rel_l2(W, G["ins_weights"]) < GOLD_TOL


def test_golden_image_shape_4x():
    G = golden("image_shape_4x")
    S = O.extract_central_slices(G["volume_rfft"], (64, 64, 64), O.prepare_rotations(G["rotations"], False, 1e-8))
    assert S.shape == G["slices"].shape == (12, 64, 33)
    assert rel_l2(S, G["slices"]) < GOLD_TOL


def test_golden_tiebreak_and_friedel():
    G = golden("tiebreak_32")
    for k in ("left", "right"):
        S = O.extract_central_slices(G["volume_rfft"], (32, 32, 32), O.prepare_rotations(G[k], False, 1e-8))
        assert rel_l2(S, G["slices_" + k]) < GOLD_TOL
    R = G["friedel_rotations"]
    S = O.extract_central_slices(G["volume_rfft"], (32, 32, 32), O.prepare_rotations(R, False, 1e-8))
    assert rel_l2(S, G["friedel_slices"]) < GOLD_TOL
    V, W = O.insert_central_slices(G["friedel_slices"], (32, 32, 32), O.prepare_rotations(R, False, None))
    assert rel_l2(V, G["friedel_volume"]) < GOLD_TOL
    assert max_abs(W, G["friedel_weights"]) < 1e-5


@pytest.mark.parametrize("mode", ["trilinear", "nearest"])
def test_golden_sample_image_3d(mode):
    G = golden("sample_image_3d")
    for key in ("real", "complex", "multi"):
        out = O.sample_image_3d(G["image_" + key], G["coordinates"], mode)
        assert out.shape == G[f"{key}_{mode}"].shape
        assert rel_l2(out, G[f"{key}_{mode}"]) < GOLD_TOL
        if mode == "nearest":  # nearest is a selection: exact, ties included
            assert max_abs(out, G[f"{key}_{mode}"]) == 0.0


@pytest.mark.parametrize("mode", ["trilinear", "nearest"])
def test_golden_insert_into_image_3d(mode):
    G = golden("insert_into_image_3d")
    for key, kk, shape in (("r", "real", (12, 10, 9)), ("c", "complex", (12, 10, 9)), ("m", "multi", (2, 12, 10, 9))):
        img = np.zeros(shape, dtype=np.complex128 if key == "c" else np.float64)
        img, w = O.insert_into_image_3d(G[f"val_{key}1"], G["coords1"], img, None, mode)
        img, w = O.insert_into_image_3d(G[f"val_{key}2"], G["coords2"], img, w, mode)
        assert rel_l2(img, G[f"{kk}_{mode}_image"]) < GOLD_TOL
        assert rel_l2(w, G[f"{kk}_{mode}_weights"]) < GOLD_TOL


def test_golden_project_backproject():
    G = golden("project_backproject_24")
    P = O.project_3d_to_2d(G["volume"], G["rotations"])
    assert rel_l2(P, G["projections"]) < GOLD_TOL
    B = O.backproject_2d_to_3d(G["projections"], G["rotations"])
    assert rel_l2(B, G["backprojection"]) < GOLD_TOL


def test_golden_autograd():
    G = golden("autograd_8")
    Rz = O.prepare_rotations(G["ext_rotations"], False, 1e-8)
    gV = O.extract_adjoint(G["ext_target"], (8, 8, 5), (8, 8, 8), Rz)
    assert rel_l2(gV, G["ext_grad_volume"]) < GOLD_TOL
    gR = O.rotation_grad_to_xyz(
        O.extract_rotation_grad(G["volume_rfft"], G["ext_target"], (8, 8, 8), Rz), G["ext_rotations"], False, 1e-8
    )
    assert rel_l2(gR, G["ext_grad_rotations"]) < GOLD_TOL
    Rz = O.prepare_rotations(G["ins_rotations"], False, None)
    gS = O.insert_adjoint(G["ins_target_volume"], (8, 8, 8), Rz)
    assert rel_l2(gS, G["ins_grad_slices"]) < GOLD_TOL
    gR = O.rotation_grad_to_xyz(
        O.insert_rotation_grad(G["ins_slices"], G["ins_target_volume"], G["ins_target_weights"], (8, 8, 8), Rz),
        G["ins_rotations"], False, None,
    )
    assert rel_l2(gR, G["ins_grad_rotations"]) < GOLD_TOL


# ------------------------------------------------------------ closed forms / invariants
def test_frequencies_closed_form():
    fy, fx = O.central_slice_frequencies(6, 6)
    assert np.allclose(fy[:, 0], np.fft.fftshift(np.fft.fftfreq(6)))
    assert np.allclose(fx[0], np.fft.rfftfreq(6))
    fy, fx = O.central_slice_frequencies(5, 5)
    assert np.allclose(fy[:, 0], np.fft.fftshift(np.fft.fftfreq(5)))
    assert np.allclose(fx[0], np.fft.rfftfreq(5))


def test_fftfreq_mask_disc():
    m = O.fftfreq_mask(16, 16, 0.25)
    fy, fx = O.central_slice_frequencies(16, 16)
    assert np.array_equal(m, np.sqrt(fy ** 2 + fx ** 2) <= 0.25 + 1e-12)
    assert O.fftfreq_mask(16, 16, None).all()


def test_extract_identity_is_central_plane():
    V = _rng_volume_rfft(16, 0)
    S = O.extract_central_slices(V, (16, 16, 16), O.prepare_rotations(np.eye(3)[None], False, 1e-8))
    assert np.array_equal(S[0], V[8])


def test_extract_rz90_is_permutation():
    """Rz(90): (x, y) -> (-y, x); pixel (fy, fx) reads V at (D/2, N/2 + N fx, -N fy)
    conjugated where fy > 0, else at (D/2, N/2 - N fx, N fy).  On the x' = 0
    row (fy = 0) R_zyx[2,2] = R[0,0] = 0, so the tie is decided by
    R_zyx[2,1] = R[0,1] = -1 < 0: those points are conjugated too."""
    N = 16
    V = _rng_volume_rfft(N, 1)
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    S = O.extract_central_slices(V, (N, N, N), O.prepare_rotations(R[None], False, 1e-8))[0]
    expect = np.zeros_like(S)
    for i in range(N):
        for j in range(N // 2 + 1):
            ky, kx = i - N // 2, j  # integer frequencies
            # rotated (x', y') = (-ky, kx); conj (and negate) where x' < 0
            xr, yr = -ky, kx
            conj = xr <= 0
            if conj:
                xr, yr = -xr, -yr
            zi, yi, xi = N // 2, yr + N // 2, xr
            if 0 <= yi < N and 0 <= xi <= N // 2:
                v = V[zi, yi, xi]
                expect[i, j] = np.conj(v) if conj else v
    assert max_abs(S, expect) < 1e-12


@pytest.mark.parametrize("seed", [0, 1])
def test_extract_conjugate_pairs(seed):
    """R and R.Rz_zyz(180) give conjugate slices (TFS tests :108-232)."""
    V = _rng_volume_rfft(24, 10 + seed)
    pairs = [
        ([0, 0, 0], [0, 0, 180]), ([0, 0, 0], [180, 0, 0]), ([0, 20, 0], [0, 20, 180]),
        ([0, 20, 90], [0, 20, 270]), ([0, 60, 0], [0, 60, 180]), ([90, 90, 0], [90, 90, 180]),
        ([45, 0, 0], [45, 0, 180]), ([45, 90, 0], [45, 90, 180]),
    ]
    left = np.stack([Rotation.from_euler("ZYZ", a, degrees=True).as_matrix() for a, _ in pairs])
    right = np.stack([Rotation.from_euler("ZYZ", b, degrees=True).as_matrix() for _, b in pairs])
    rnd = _rotations(20, seed)
    off = Rotation.from_euler("ZYZ", [0, 0, 180], degrees=True).as_matrix()
    left = np.concatenate([left, rnd])
    right = np.concatenate([right, rnd @ off])
    SL = O.extract_central_slices(V, (24, 24, 24), O.prepare_rotations(left, False, 1e-8))
    SR = O.extract_central_slices(V, (24, 24, 24), O.prepare_rotations(right, False, 1e-8))
    assert max_abs(SL, np.conj(SR)) < 1e-6 * np.abs(V).max()


def test_extract_x0_line_friedel():
    V = _rng_volume_rfft(24, 3)
    S = O.extract_central_slices(V, (24, 24, 24), O.prepare_rotations(_rotations(10, 3), False, 1e-8))
    line = S[:, :, 0]
    half = line[:, 1:12]
    mate = np.conj(line[:, -11:])[:, ::-1]
    assert max_abs(half, mate) < 1e-9 * np.abs(V).max()


def test_extract_adjoint_dot_product():
    rng = np.random.default_rng(4)
    V = rng.standard_normal((20, 20, 11)) + 1j * rng.standard_normal((20, 20, 11))
    Rz = O.prepare_rotations(_rotations(7, 4), False, 1e-8)
    T = rng.standard_normal((7, 20, 11)) + 1j * rng.standard_normal((7, 20, 11))
    lhs = np.real(np.vdot(T, O.extract_central_slices(V, (20, 20, 20), Rz)))
    rhs = np.real(np.vdot(O.extract_adjoint(T, V.shape, (20, 20, 20), Rz), V))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_insert_adjoint_dot_product():
    rng = np.random.default_rng(5)
    d = rng.standard_normal((6, 18, 10)) + 1j * rng.standard_normal((6, 18, 10))
    Rz = O.prepare_rotations(_rotations(6, 5), False, None)
    T = rng.standard_normal((18, 18, 10)) + 1j * rng.standard_normal((18, 18, 10))
    V, _ = O.insert_central_slices(d, (18, 18, 18), Rz)
    lhs = np.real(np.vdot(T, V))
    rhs = np.real(np.vdot(O.insert_adjoint(T, (18, 18, 18), Rz), d))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def _fd_rotation_grad(fn, Rz, h=1e-7):
    g = np.zeros_like(Rz)
    for idx in np.ndindex(*Rz.shape):
        Rp, Rm = Rz.copy(), Rz.copy()
        Rp[idx] += h
        Rm[idx] -= h
        g[idx] = (fn(Rp) - fn(Rm)) / (2 * h)
    return g


def test_extract_rotation_grad_finite_differences():
    rng = np.random.default_rng(6)
    V = O.volume_rfft(rng.standard_normal((12, 12, 12)))
    Rz = O.prepare_rotations(_rotations(3, 6), False, 1e-8).astype(np.float64)
    T = rng.standard_normal((3, 12, 7)) + 1j * rng.standard_normal((3, 12, 7))
    loss = lambda R: np.real(np.vdot(T, O.extract_central_slices(V, (12, 12, 12), R)))
    g = O.extract_rotation_grad(V, T, (12, 12, 12), Rz)
    fd = _fd_rotation_grad(loss, Rz)
    assert rel_l2(g, fd) < 1e-5
    assert np.all(g[:, :, 0] == 0)  # the z=0 slice plane has no z column gradient


def test_insert_rotation_grad_finite_differences():
    rng = np.random.default_rng(7)
    d = rng.standard_normal((3, 12, 7)) + 1j * rng.standard_normal((3, 12, 7))
    Rz = O.prepare_rotations(_rotations(3, 7), False, None).astype(np.float64)
    TV = rng.standard_normal((12, 12, 7)) + 1j * rng.standard_normal((12, 12, 7))
    TW = rng.standard_normal((12, 12, 7))

    def loss(R):
        V, W = O.insert_central_slices(d, (12, 12, 12), R)
        return np.real(np.vdot(TV, V)) + np.sum(TW * W)

    g = O.insert_rotation_grad(d, TV, TW, (12, 12, 12), Rz)
    fd = _fd_rotation_grad(loss, Rz)
    assert rel_l2(g, fd) < 1e-5


def test_insert_identity_known_answer():
    """R = I: every pixel lands on voxel (D/2, i, j) with weight 1; the x=0
    column keeps fy >= 0 pixels and mirrors them to (D/2, H-i, 0) conjugated;
    DC is not mirrored; (D/2, 0, 0) receives nothing."""
    N = 8
    rng = np.random.default_rng(8)
    d = rng.standard_normal((1, N, N // 2 + 1)) + 1j * rng.standard_normal((1, N, N // 2 + 1))
    V, W = O.insert_central_slices(d, (N, N, N), O.prepare_rotations(np.eye(3)[None], False, None))
    eV = np.zeros_like(V)
    eW = np.zeros_like(W)
    eV[N // 2, :, 1:] = d[0, :, 1:]
    eW[N // 2, :, 1:] = 1
    for i in range(N // 2, N):
        eV[N // 2, i, 0] = d[0, i, 0]
        eW[N // 2, i, 0] = 1
    for i in range(N // 2 + 1, N):
        eV[N // 2, N - i, 0] = np.conj(d[0, i, 0])
        eW[N // 2, N - i, 0] = 1
    assert max_abs(V, eV) < 1e-12 and max_abs(W, eW) < 1e-12


def test_insert_friedel_x0_plane():
    """x=0 plane of an insertion obeys F(0,y,z) = conj F(0,-y,-z) (TFS tests :188-301)."""
    N = 16
    rng = np.random.default_rng(9)
    vol = rng.standard_normal((N, N, N))
    R = _rotations(8, 9)
    Rz = O.prepare_rotations(R, False, 1e-8)
    S = O.extract_central_slices(O.volume_rfft(vol), (N, N, N), Rz)
    V, _ = O.insert_central_slices(S, (N, N, N), O.prepare_rotations(R, False, None))
    x0 = V[:, :, 0]
    zz, yy = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    mate = np.conj(x0[(N - zz) % N, (N - yy) % N])
    err = np.abs(x0 - mate)
    err[N // 2, N // 2] = 0  # DC is not mirrored (the reference test skips it too)
    assert err.max() < 1e-9 * np.abs(x0).max()


def test_sample_trilinear_reproduces_affine():
    rng = np.random.default_rng(10)
    z, y, x = np.meshgrid(np.arange(7.0), np.arange(9.0), np.arange(5.0), indexing="ij")
    img = 0.3 * z - 1.7 * y + 2.1 * x + 0.5
    c = rng.uniform(0, 1, (500, 3)) * np.array([6, 8, 4])
    expect = 0.3 * c[:, 0] - 1.7 * c[:, 1] + 2.1 * c[:, 2] + 0.5
    assert max_abs(O.sample_image_3d(img, c), expect) < 1e-12
    out = O.sample_image_3d(img, np.array([[-0.1, 1, 1], [1, 8.01, 1], [6, 8, 4]]))
    assert out[0] == 0 and out[1] == 0 and abs(out[2] - (1.8 - 13.6 + 8.4 + 0.5)) < 1e-12


def test_splat_known_answers():
    """TII tests :66-81 and :100-113."""
    img = np.zeros((28, 28, 28))
    img, w = O.insert_into_image_3d(np.array([5.0]), np.array([[10.5, 14.5, 18.5]]), img)
    assert np.allclose(img[10:12, 14:16, 18:20], 5 / 8) and np.isclose(img.sum(), 5)
    assert np.isclose(w.sum(), 1)
    img = np.zeros((28, 28, 28))
    img, _ = O.insert_into_image_3d(np.array([5.0]), np.array([[10.7, 14.3, 18.9]]), img, None, "nearest")
    e = np.zeros((28, 28, 28))
    e[11, 14, 19] = 5
    assert np.array_equal(img, e)
    img = np.zeros((4, 4, 4))
    img, w = O.insert_into_image_3d(np.array([1.0]), np.array([[1.0, 1.0, 1.0]]), img, None, "bilinear")
    assert not img.any() and not w.any()  # unknown mode: no-op


def test_project_identity_is_z_sum():
    """Central-slice theorem at R = I: the projection is the z-sum of the padded,
    sinc^2-weighted volume, cropped back."""
    rng = np.random.default_rng(11)
    vol = rng.standard_normal((12, 12, 12))
    P = O.project_3d_to_2d(vol, np.eye(3)[None])[0]
    padded = np.pad(vol, 6) * O._sinc2_grid((24, 24, 24))
    expect = padded.sum(axis=0)[6:-6, 6:-6]
    assert max_abs(P, expect) < 1e-10


def test_project_rotation_centre():
    vol = np.zeros((16, 16, 16))
    vol[8, 8, 8] = 1
    P = O.project_3d_to_2d(vol, _rotations(10, 12))
    for img in P:
        assert np.unravel_index(np.argmax(img), img.shape) == (8, 8)


def test_project_errors():
    with pytest.raises(ValueError):
        O.project_3d_to_2d(np.zeros((4, 4, 5)), np.eye(3)[None])
    with pytest.raises(ValueError):
        O.backproject_2d_to_3d(np.zeros((1, 4, 5)), np.eye(3)[None])


# ------------------------------------------------------------- the C restatement
@pytest.fixture(scope="module")
def C():
    from oracle import c_oracle

    c_oracle.build()
    return c_oracle


@pytest.mark.parametrize("name,shape,fmax", [
    ("extract_insert_64", (64, 64, 64), None), ("odd_33", (33, 33, 33), None), ("fftfreq_max_03", (32, 32, 32), 0.3),
])
def test_c_oracle_matches_numpy_oracle(C, name, shape, fmax):
    G = golden(name)
    Rz = O.prepare_rotations(G["rotations"], False, 1e-8)
    S = C.extract(G["volume_rfft"], shape, Rz, fmax)
    assert rel_l2(S, O.extract_central_slices(G["volume_rfft"], shape, Rz, fmax)) < 1e-14
    assert rel_l2(S, G["slices"]) < GOLD_TOL
    gA = C.extract_adjoint(G["slices"], G["volume_rfft"].shape, shape, Rz, fmax)
    assert rel_l2(gA, O.extract_adjoint(G["slices"], G["volume_rfft"].shape, shape, Rz, fmax)) < 1e-14
    Ri = O.prepare_rotations(G["rotations"], False, None)
    V, W = C.insert(G["slices"], shape, Ri, fmax)
    rV, rW = O.insert_central_slices(G["slices"], shape, Ri, fmax)
    assert rel_l2(V, rV) < 1e-14 and rel_l2(W, rW) < 1e-14
    assert rel_l2(V, G["ins_volume"]) < GOLD_TOL


def test_c_oracle_image_shape_and_friedel(C):
    G = golden("image_shape_4x")
    S = C.extract(G["volume_rfft"], (64, 64, 64), O.prepare_rotations(G["rotations"], False, 1e-8))
    assert rel_l2(S, G["slices"]) < GOLD_TOL
    G = golden("tiebreak_32")
    V, W = C.insert(G["friedel_slices"], (32, 32, 32), O.prepare_rotations(G["friedel_rotations"], False, None))
    assert rel_l2(V, G["friedel_volume"]) < GOLD_TOL and max_abs(W, G["friedel_weights"]) < 1e-5
