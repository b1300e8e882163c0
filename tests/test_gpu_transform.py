"""GPU parity of the fused affine warps (torch-transform-image drop-ins)
against the golden fixtures written from the reference and the FP64 oracle."""

import numpy as np
import pytest
import torch

import teamtomo_amd as tt
from conftest import golden, max_abs, rel_l2
from oracle import transform_image as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("key", ["real", "complex", "multi"])
def test_affine_3d_vs_golden(device, key):
    G = golden("transform_image")
    img, M = _t(G[f"img3_{key}"], device), _t(G["M3"], device)
    for mode in ("trilinear", "nearest"):
        out = _np(tt.affine_transform_image_3d(img, M, mode))
        assert out.shape == G[f"a3_{key}_{mode}"].shape
        assert rel_l2(out, G[f"a3_{key}_{mode}"]) < TOL
        out = _np(tt.affine_transform_image_3d(img, M, mode, output_shape=(8, 7, 6), zyx_matrices=True))
        assert rel_l2(out, G[f"a3_{key}_{mode}_zyx_out"]) < TOL
        if mode == "nearest":
            assert max_abs(out, G[f"a3_{key}_{mode}_zyx_out"]) == 0.0
    single = tt.affine_transform_image_3d(_t(G["img3_real"], device), M[1, 0, 0, 0], "trilinear")
    assert rel_l2(_np(single), G["a3_single"]) < TOL


@pytest.mark.parametrize("key", ["real", "complex", "multi"])
def test_affine_2d_vs_golden(device, key):
    G = golden("transform_image")
    img, M = _t(G[f"img2_{key}"], device), _t(G["M2"], device)
    for mode in ("bilinear", "nearest", "bicubic"):
        out = _np(tt.affine_transform_image_2d(img, M, mode))
        assert out.shape == G[f"a2_{key}_{mode}"].shape
        assert rel_l2(out, G[f"a2_{key}_{mode}"]) < TOL
        out = _np(tt.affine_transform_image_2d(img, M, mode, output_shape=(9, 8), yx_matrices=True))
        assert rel_l2(out, G[f"a2_{key}_{mode}_yx_out"]) < TOL


def test_rotate_shift_and_per_column_vs_golden(device):
    G = golden("transform_image")
    img = _t(G["img2_real"], device)
    for mode in ("bilinear", "bicubic"):
        assert rel_l2(_np(tt.rotate_then_shift_image_2d(img, 30.0, (2.5, -1.5), mode)), G[f"rts_{mode}"]) < TOL
        assert rel_l2(_np(tt.shift_then_rotate_image_2d(img, 30.0, (2.5, -1.5), mode)), G[f"str_{mode}"]) < TOL
    assert rel_l2(_np(tt.rotate_then_shift_image_2d(img, shift=(1.0, 2.0), interpolation="bilinear")),
                  G["shift_only"]) < TOL
    assert rel_l2(_np(tt.affine_transform_image_2d(img, _t(G["Mcol"], device), "bilinear")), G["a2_percol"]) < TOL


def test_affine_3d_large_vs_oracle_and_errors(device):
    # a 128^3 volume warped by 6 matrices (exercises the grid-stride walk)
    rng = np.random.default_rng(3)
    img = rng.standard_normal((64, 80, 96)).astype(np.float32)
    from scipy.spatial.transform import Rotation
    M = np.tile(np.eye(4), (6, 1, 1))
    c = (np.array([96, 80, 64]) - 1) / 2
    for k in range(6):
        A = Rotation.random(random_state=k).as_matrix()
        M[k, :3, :3], M[k, :3, 3] = A, c - A @ c + rng.uniform(-3, 3, 3)
    M = M.astype(np.float32).reshape(6, 1, 1, 1, 4, 4)
    out = _np(tt.affine_transform_image_3d(_t(img, device), _t(M, device), "trilinear", output_shape=(40, 50, 60)))
    ref = O.affine_transform_image_3d(img, M, "trilinear", output_shape=(40, 50, 60))
    assert out.shape == (6, 40, 50, 60)
    assert rel_l2(out, ref) < TOL
    with pytest.raises(RuntimeError):  # float64 matrices: the reference's matmul dtype error
        tt.affine_transform_image_3d(_t(img, device), torch.eye(4, dtype=torch.float64, device=device), "trilinear")
    with pytest.raises(RuntimeError):  # float64 image: grid_sample's dtype error
        tt.affine_transform_image_3d(_t(img, device).double(), torch.eye(4, device=device), "trilinear")
    with pytest.raises(ValueError):
        tt.affine_transform_image_3d(_t(img, device), torch.eye(4, device=device), "bicubic")
    with pytest.raises(RuntimeError):  # matrices that do not broadcast against the grid
        tt.affine_transform_image_3d(_t(img, device), torch.eye(4, device=device).expand(7, 4, 4), "trilinear")
