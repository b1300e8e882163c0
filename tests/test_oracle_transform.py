"""The FP64 oracle of the affine warps (torch-transform-image), pinned against
the golden fixtures written from the reference and against exact cases: the
identity, integer translations and 90-degree rotations about the centre."""

import numpy as np
import pytest

from conftest import golden, max_abs, rel_l2
from oracle import transform_image as O

TOL = 2e-6


@pytest.mark.parametrize("key", ["real", "complex", "multi"])
def test_golden_affine_3d(key):
    G = golden("transform_image")
    img = G[f"img3_{key}"]
    for mode in ("trilinear", "nearest"):
        out = O.affine_transform_image_3d(img, G["M3"], mode)
        assert out.shape == G[f"a3_{key}_{mode}"].shape
        assert rel_l2(out, G[f"a3_{key}_{mode}"]) < TOL
        out = O.affine_transform_image_3d(img, G["M3"], mode, output_shape=(8, 7, 6), zyx_matrices=True)
        assert rel_l2(out, G[f"a3_{key}_{mode}_zyx_out"]) < TOL
    assert rel_l2(O.affine_transform_image_3d(G["img3_real"], G["M3"][1, 0, 0, 0], "trilinear"), G["a3_single"]) < TOL


@pytest.mark.parametrize("key", ["real", "complex", "multi"])
def test_golden_affine_2d(key):
    G = golden("transform_image")
    img = G[f"img2_{key}"]
    for mode in ("bilinear", "nearest", "bicubic"):
        assert rel_l2(O.affine_transform_image_2d(img, G["M2"], mode), G[f"a2_{key}_{mode}"]) < TOL
        out = O.affine_transform_image_2d(img, G["M2"], mode, output_shape=(9, 8), yx_matrices=True)
        assert rel_l2(out, G[f"a2_{key}_{mode}_yx_out"]) < TOL


def test_golden_rotate_shift_and_per_column_matrices():
    G = golden("transform_image")
    img = G["img2_real"]
    for mode in ("bilinear", "bicubic"):
        assert rel_l2(O.rotate_then_shift_image_2d(img, 30.0, (2.5, -1.5), mode), G[f"rts_{mode}"]) < TOL
        assert rel_l2(O.shift_then_rotate_image_2d(img, 30.0, (2.5, -1.5), mode), G[f"str_{mode}"]) < TOL
    assert rel_l2(O.rotate_then_shift_image_2d(img, shift=(1.0, 2.0), interpolation="bilinear"),
                  G["shift_only"]) < TOL
    assert rel_l2(O.affine_transform_image_2d(img, G["Mcol"], "bilinear"), G["a2_percol"]) < TOL


def test_exact_cases():
    rng = np.random.default_rng(0)
    img = rng.standard_normal((6, 7, 8))
    for mode in ("trilinear", "nearest"):
        assert max_abs(O.affine_transform_image_3d(img, np.eye(4), mode), img) == 0.0
    # xyz translation by (+1, 0, -2): output(z, y, x) = img(z - 2, y, x + 1)
    M = np.eye(4)
    M[:3, 3] = [1, 0, -2]
    out = O.affine_transform_image_3d(img, M, "trilinear")
    assert max_abs(out[2:, :, :-1], img[:-2, :, 1:]) == 0.0
    assert np.all(out[:2] == 0) and np.all(out[:, :, -1] == 0)
    # 2-D: a 90-degree rotation about the centre of a square image is a permutation
    sq = rng.standard_normal((9, 9))
    R90 = np.array([[0, -1, 8], [1, 0, 0], [0, 0, 1]], dtype=np.float64)  # yx: (y, x) -> (8 - x, y)
    out = O.affine_transform_image_2d(sq, R90, "bilinear", yx_matrices=True)
    assert max_abs(out, sq[8 - np.arange(9)[None, :], np.arange(9)[:, None]]) < 1e-12
