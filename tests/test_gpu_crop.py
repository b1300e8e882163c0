"""GPU parity of the subpixel crops (torch-subpixel-crop drop-ins) against the
golden fixture written from the reference and the FP64 oracle."""

import numpy as np
import pytest
import torch

import teamtomo_amd as tt
from conftest import golden, rel_l2
from oracle import subpixel_crop as O

pytestmark = pytest.mark.gpu


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def _np(t):
    return t.detach().cpu().numpy()


def test_crop_3d_vs_golden(device):
    G = golden("subpixel_crop")
    vol, pos = _t(G["vol"], device), _t(G["pos3"], device)
    assert rel_l2(_np(tt.subpixel_crop_3d(vol, pos, 8)), G["c3"]) < 1e-5
    assert rel_l2(_np(tt.subpixel_crop_3d(vol, pos, 8, mask=_t(G["mask3"], device))), G["c3_mask"]) < 1e-5
    out = tt.subpixel_crop_3d(vol, pos, 8, return_rfft=True)
    assert out.dtype == torch.complex64 and rel_l2(_np(out), G["c3_rfft"]) < 1e-5
    assert rel_l2(_np(tt.subpixel_crop_3d(vol, pos, 8, return_rfft=True, decenter=True)), G["c3_rfft_dec"]) < 1e-5
    # a mask that only broadcasts is applied after the crop
    m = _t(G["mask3"], device)[None]
    assert rel_l2(_np(tt.subpixel_crop_3d(vol, pos, 8, mask=m)), G["c3_mask"]) < 1e-5


def test_crop_2d_vs_golden(device):
    G = golden("subpixel_crop")
    imgs, pos = _t(G["imgs"], device), _t(G["pos2"], device)
    assert rel_l2(_np(tt.subpixel_crop_2d(imgs, pos, 10)), G["c2"]) < 1e-5
    assert rel_l2(_np(tt.subpixel_crop_2d(imgs, pos, 10, mask=_t(G["mask2"], device))), G["c2_mask"]) < 1e-5
    assert rel_l2(_np(tt.subpixel_crop_2d(imgs, pos, 10, return_rfft=True, decenter=True)), G["c2_rfft_dec"]) < 1e-5
    single = _np(tt.subpixel_crop_2d(imgs[1], pos[:, 1], 9))
    assert single.shape == G["c2_single"].shape and rel_l2(single, G["c2_single"]) < 1e-5


def test_crop_3d_many_patches_vs_oracle(device):
    rng = np.random.default_rng(5)
    vol = rng.standard_normal((64, 60, 56)).astype(np.float32)
    pos = (rng.uniform(-4, 1, (3, 50, 3)) + np.array([32, 30, 28])).astype(np.float32)
    pos[0, :5] = rng.uniform(-3, 66, (5, 3))  # patches over the edges
    out = _np(tt.subpixel_crop_3d(_t(vol, device), _t(pos, device), 16))
    ref = O.subpixel_crop_3d(vol, pos, 16)
    assert out.shape == (3, 50, 16, 16, 16)
    assert rel_l2(out, ref) < 1e-5
    with pytest.raises(RuntimeError):
        tt.subpixel_crop_3d(_t(vol, device).double(), _t(pos, device), 16)
    with pytest.raises(ValueError):
        tt.subpixel_crop_2d(torch.zeros((2, 8, 8), device=device), torch.zeros((4, 3, 2), device=device), 4)


def test_crop_without_positions(device):
    """No positions: empty patches (real or rfft) rather than an FFT error."""
    vol = torch.rand(16, 16, 16, device=device)
    assert tt.subpixel_crop_3d(vol, torch.zeros(0, 3, device=device), 8).shape == (0, 8, 8, 8)
    out = tt.subpixel_crop_3d(vol, torch.zeros(0, 3, device=device), 8, return_rfft=True)
    assert out.shape == (0, 8, 8, 5) and out.dtype == torch.complex64
    imgs = torch.rand(2, 16, 16, device=device)
    assert tt.subpixel_crop_2d(imgs, torch.zeros(0, 2, 2, device=device), 8).shape == (0, 2, 8, 8)
