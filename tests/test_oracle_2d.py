"""The FP64 oracle of the 2-D analogues, pinned against the reference's golden
fixtures (written from torch-fourier-slice / torch-image-interpolation by
tests/golden/make_golden.py) and against what the mathematics fixes: the R = I
line, exact 90-degree permutations, the x' < 0 conjugate rule, adjointness,
affine reproduction (bilinear), constants and the Keys quadratic identity
(bicubic weights), and the reference tests' known answers."""

import numpy as np
import pytest

from conftest import golden, max_abs, rel_l2
from oracle import slice_2d as O

TOL = 2e-6  # float32 reference vs FP64 oracle


def _rot(theta):
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s], [s, c]])


@pytest.mark.parametrize("key,shape", [("even", (32, 32)), ("odd", (33, 33)), ("shape2x", (32, 32))])
def test_golden_lines(key, shape):
    G = golden("extract_2d")
    F, R = G[f"{key}_image_rfft"], G[f"{key}_rotations"]
    assert rel_l2(O.extract_central_lines(F, shape, O.prepare_rotations_2d(R)), G[f"{key}_lines"]) < TOL
    assert rel_l2(O.extract_central_lines(F, shape, O.prepare_rotations_2d(R, True)),
                  G[f"{key}_lines_yx"]) < TOL


def test_golden_project_2d_to_1d():
    G = golden("extract_2d")
    assert rel_l2(O.project_2d_to_1d(G["proj_image"], G["proj_rotations"]), G["proj_pad2"]) < TOL
    assert rel_l2(O.project_2d_to_1d(G["proj_image"], G["proj_rotations"], pad_factor=1.0),
                  G["proj_pad1"]) < TOL


@pytest.mark.parametrize("mode", ["bilinear", "nearest", "bicubic"])
def test_golden_sample_image_2d(mode):
    G = golden("sample_image_2d")
    for key in ("real", "complex", "multi"):
        out = O.sample_image_2d(G[f"image_{key}"], G["coordinates"], mode)
        assert rel_l2(out, G[f"{key}_{mode}"]) < TOL
        if mode == "nearest":
            assert max_abs(out, G[f"{key}_{mode}"]) == 0.0


@pytest.mark.parametrize("mode", ["bilinear", "nearest"])
def test_golden_insert_into_image_2d(mode):
    G = golden("insert_into_image_2d")
    for key, v1, v2, shape, dt in (("real", "val_r1", "val_r2", (12, 10), np.float64),
                                   ("complex", "val_c1", "val_c2", (12, 10), np.complex128),
                                   ("multi", "val_m1", "val_m2", (2, 12, 10), np.float64)):
        img, w = O.insert_into_image_2d(G[v1], G["coords1"], np.zeros(shape, dt), None, mode)
        img, w = O.insert_into_image_2d(G[v2], G["coords2"], img, w, mode)
        assert rel_l2(img, G[f"{key}_{mode}_image"]) < 1e-6
        assert rel_l2(w, G[f"{key}_{mode}_weights"]) < 1e-6


def test_identity_line_is_central_row():
    rng = np.random.default_rng(0)
    F = rng.standard_normal((16, 9)) + 1j * rng.standard_normal((16, 9))
    L = O.extract_central_lines(F, (16, 16), O.prepare_rotations_2d(np.eye(2)[None]))
    assert max_abs(L[0], F[8]) == 0.0


def test_rotation_90_reads_the_column():
    # xy rotation by +90 degrees takes (x, y) = (f, 0) to (0, f): the line runs
    # up the y axis from DC, i.e. down the x = 0 column from row H//2
    rng = np.random.default_rng(1)
    F = rng.standard_normal((16, 9)) + 1j * rng.standard_normal((16, 9))
    R = np.array([[[0.0, -1.0], [1.0, 0.0]]])
    L = O.extract_central_lines(F, (16, 16), O.prepare_rotations_2d(R))
    j = np.arange(9)
    ok = 8 + j <= 15
    assert max_abs(L[0][ok], F[8 + j[ok], 0]) == 0.0
    assert np.all(L[0][~ok] == 0)


def test_half_turn_reads_conjugates():
    # a 180-degree rotation sends (f, 0) to (-f, 0): read at (H//2, f), conjugated
    rng = np.random.default_rng(2)
    F = rng.standard_normal((16, 9)) + 1j * rng.standard_normal((16, 9))
    L0 = O.extract_central_lines(F, (16, 16), O.prepare_rotations_2d(np.eye(2)[None]))
    L1 = O.extract_central_lines(F, (16, 16), O.prepare_rotations_2d(-np.eye(2)[None]))
    assert max_abs(L1[0][1:], np.conj(L0[0][1:])) < 1e-12
    assert L1[0][0] == L0[0][0]  # DC: x' = 0 is not conjugated


def test_line_adjoint_dot_product():
    rng = np.random.default_rng(3)
    shape = (20, 11)
    F = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    R = O.prepare_rotations_2d(np.stack([_rot(t) for t in rng.uniform(0, 2 * np.pi, 7)]))
    G = rng.standard_normal((7, 11)) + 1j * rng.standard_normal((7, 11))
    lhs = np.real(np.vdot(G, O.extract_central_lines(F, (20, 20), R)))
    rhs = np.real(np.vdot(O.extract_central_lines_adjoint(G, shape, (20, 20), R), F))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_samplers_reproduce_affine_functions():
    rng = np.random.default_rng(4)
    yy, xx = np.meshgrid(np.arange(12.0), np.arange(10.0), indexing="ij")
    img = 0.7 * yy - 1.3 * xx + 2.0
    c = np.stack([rng.uniform(0, 11, 300), rng.uniform(0, 9, 300)], -1)
    exact = 0.7 * c[:, 0] - 1.3 * c[:, 1] + 2.0
    assert max_abs(O.sample_image_2d(img, c, "bilinear"), exact) < 1e-12
    # grid_sample's cubic convolution (Keys, A = -0.75) reproduces constants
    # wherever all 16 taps are inside the image (clipped border taps aside)...
    inner = np.all((c >= 1) & (c <= np.array([12 - 3, 10 - 3])), axis=-1)
    assert max_abs(O.sample_image_2d(np.full((12, 10), 2.5), c[inner], "bicubic"), 2.5) < 1e-12
    # ...and the same weight formula with A = -0.5 reproduces quadratics (Keys 1981),
    # which a wrong tap order or sign would break
    t = rng.uniform(0, 1, 200)
    w = O._cubic_weights(t, A=-0.5)
    for poly in (lambda x: x, lambda x: x * x - 3 * x + 1):
        est = sum(w[a] * poly(a - 1.0) for a in range(4))
        assert max_abs(est, poly(t)) < 1e-12
    # and interpolates: integer coordinates return the pixel
    ints = np.stack([rng.integers(0, 12, 50), rng.integers(0, 10, 50)], -1).astype(float)
    for mode in ("bilinear", "nearest", "bicubic"):
        assert max_abs(O.sample_image_2d(img, ints, mode), img[ints[:, 0].astype(int), ints[:, 1].astype(int)]) < 1e-12


def test_samplers_zero_outside():
    img = np.ones((5, 4))
    c = np.array([[-0.01, 1.0], [1.0, 3.01], [4.0, 3.0], [0.0, 0.0]])
    for mode in ("bilinear", "nearest", "bicubic"):
        assert np.allclose(O.sample_image_2d(img, c, mode), [0, 0, 1, 1])


def test_splat_known_answers():
    # TII tests/test_image_interpolation_2d.py:65-80 and :143-160
    img, _ = O.insert_into_image_2d(np.array([5.0]), np.array([[10.5, 14.5]]), np.zeros((28, 28)))
    assert np.allclose(img[10:12, 14:16], 1.25) and img.sum() == pytest.approx(5.0)
    img, w = O.insert_into_image_2d(np.array([5.0]), np.array([[10.7, 14.3]]), np.zeros((28, 28)),
                                    interpolation="nearest")
    expected = np.zeros((28, 28))
    expected[11, 14] = 5
    assert np.array_equal(img, expected) and w[11, 14] == 1.0


def test_project_2d_rotation_centre():
    # TFS tests/test_torch_fourier_slice.py:53-76: the DC pixel never moves
    img = np.zeros((32, 32))
    img[16, 16] = 1
    rng = np.random.default_rng(5)
    R = np.stack([_rot(t) for t in rng.uniform(0, 2 * np.pi, 20)])
    P = O.project_2d_to_1d(img, R)
    assert np.all(np.argmax(P, axis=-1) == 16)


def test_project_2d_identity_is_column_sum():
    # central-slice theorem: with R = I the line (0, f_x) is the 1-D DFT of the
    # sum over y of the sinc^2-corrected padded image
    rng = np.random.default_rng(6)
    img = rng.standard_normal((16, 16))
    P = O.project_2d_to_1d(img, np.eye(2)[None], pad_factor=1.0)
    ref = (img * O._sinc2_grid_2d((16, 16))).sum(axis=0)
    assert max_abs(P[0], ref) < 1e-12


# ---------------------------------------------------------------- 1-D samplers


def test_1d_splat_known_answers():
    """TII tests/test_image_interpolation_1d.py:66-160: 1 at 14.28 splits
    0.72 / 0.28 over pixels 14 and 15; 5 at 10.7 (nearest) lands on 11."""
    img, w = O.insert_into_image_1d(np.array([1.0]), np.array([14.28]), np.zeros(28), interpolation="linear")
    want = np.zeros(28)
    want[14], want[15] = 0.72, 0.28
    np.testing.assert_allclose(img, want, atol=1e-12)
    np.testing.assert_allclose(w, want, atol=1e-12)
    img, _ = O.insert_into_image_1d(np.array([5.0]), np.array([10.7]), np.zeros(28), interpolation="nearest")
    want = np.zeros(28)
    want[11] = 5.0
    np.testing.assert_allclose(img, want)
    # an unknown mode changes nothing; points outside [0, w-1] are dropped
    img, w = O.insert_into_image_1d(np.ones(3), np.array([1.5, -0.1, 27.01]), np.zeros(28), interpolation="cubic")
    assert not img.any() and not w.any()
    img, w = O.insert_into_image_1d(np.ones(3), np.array([1.5, -0.1, 27.01]), np.zeros(28))
    assert img.sum() == pytest.approx(1.0) and w.sum() == pytest.approx(1.0)


def test_1d_samplers_reproduce_lines_and_constants():
    rng = np.random.default_rng(4)
    w = 30
    x = np.arange(w, dtype=np.float64)
    line = 0.7 * x - 3.0
    c = rng.uniform(0, w - 1, 500)
    np.testing.assert_allclose(O.sample_image_1d(line, c, "linear"), 0.7 * c - 3.0, atol=1e-12)
    np.testing.assert_allclose(O.sample_image_1d(np.full(w, 2.5), c, "cubic"), 2.5, atol=1e-12)
    ints = rng.integers(0, w, 50).astype(np.float64)
    img = rng.standard_normal(w)
    for mode in ("linear", "cubic", "nearest"):
        np.testing.assert_allclose(O.sample_image_1d(img, ints, mode), img[ints.astype(int)], atol=1e-12)
        assert not O.sample_image_1d(img, np.array([-0.01, w - 0.99]), mode).any()


def test_1d_sampler_is_the_2d_sampler_on_the_repeated_image():
    """The reference's 1-D sampler is grid_sample on the image repeated to
    height 2 at y = 0; the 2-D oracle is pinned against its goldens."""
    rng = np.random.default_rng(5)
    img = rng.standard_normal((3, 25)) + 1j * rng.standard_normal((3, 25))
    c = rng.uniform(-2, 26, (4, 60))
    c[0, :5] = [0.0, 24.0, 12.5, 3.5, 7.5]  # ties for nearest
    img2 = np.repeat(img[:, None, :], 2, axis=1)
    c2 = np.stack([np.zeros_like(c), c], -1)
    for m1, m2 in (("linear", "bilinear"), ("cubic", "bicubic"), ("nearest", "nearest")):
        np.testing.assert_allclose(O.sample_image_1d(img, c, m1), O.sample_image_2d(img2, c2, m2), atol=1e-12)
