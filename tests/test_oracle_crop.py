"""The FP64 oracle of torch-subpixel-crop, pinned against the golden fixture
written from the reference, integer positions (a plain crop) and the shift
theorem."""

import numpy as np

from conftest import golden, max_abs, rel_l2
from oracle import subpixel_crop as O

TOL = 2e-6


def test_golden_crop_3d():
    G = golden("subpixel_crop")
    assert rel_l2(O.subpixel_crop_3d(G["vol"], G["pos3"], 8), G["c3"]) < TOL
    assert rel_l2(O.subpixel_crop_3d(G["vol"], G["pos3"], 8, mask=G["mask3"]), G["c3_mask"]) < TOL
    assert rel_l2(O.subpixel_crop_3d(G["vol"], G["pos3"], 8, return_rfft=True), G["c3_rfft"]) < TOL
    assert rel_l2(O.subpixel_crop_3d(G["vol"], G["pos3"], 8, return_rfft=True, decenter=True),
                  G["c3_rfft_dec"]) < 1e-5


def test_golden_crop_2d():
    G = golden("subpixel_crop")
    assert rel_l2(O.subpixel_crop_2d(G["imgs"], G["pos2"], 10), G["c2"]) < TOL
    assert rel_l2(O.subpixel_crop_2d(G["imgs"], G["pos2"], 10, mask=G["mask2"]), G["c2_mask"]) < TOL
    assert rel_l2(O.subpixel_crop_2d(G["imgs"], G["pos2"], 10, return_rfft=True, decenter=True),
                  G["c2_rfft_dec"]) < 1e-5
    single = O.subpixel_crop_2d(G["imgs"][1], G["pos2"][:, 1], 9)
    assert single.shape == G["c2_single"].shape and rel_l2(single, G["c2_single"]) < TOL


def test_integer_positions_are_plain_crops():
    rng = np.random.default_rng(0)
    vol = rng.standard_normal((12, 11, 10))
    out = O.subpixel_crop_3d(vol, np.array([[5.0, 6.0, 4.0]]), 4)[0]
    assert max_abs(out, vol[3:7, 4:8, 2:6]) < 1e-12
    img = rng.standard_normal((9, 8))
    out = O.subpixel_crop_2d(img, np.array([1.0, 1.0]), 4)  # crosses the corner: zeros outside
    ref = np.zeros((4, 4))
    ref[1:, 1:] = img[:3, :3]
    assert max_abs(out, ref) < 1e-12


def test_shift_theorem():
    # a band-limited periodic patch moved by half a pixel equals the patch
    # evaluated half a pixel away
    n = 16
    y, x = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    f = lambda yy, xx: np.cos(2 * np.pi * 2 * yy / n) + np.sin(2 * np.pi * 3 * xx / n)
    dft = np.fft.rfftn(f(y, x))
    moved = np.real(np.fft.irfftn(O.phase_shift_rfft(dft, np.array([[0.5, -0.25]]), (n, n))[0]))
    assert max_abs(moved, f(y - 0.5, x + 0.25)) < 1e-12
