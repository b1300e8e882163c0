"""The binned scatter (pixels sorted into per-brick bins, one LDS pass per
brick) against the FP64 C oracle, for the insertion and the extraction
adjoint: tilt series with an in-plane angle (the workload it is chosen for),
exact-geometry rotations whose pixels land on voxels and brick faces, many
rotations (large bins and wide sums), the disc mask, multichannel insertion
with shared weights, empty calls, and the choice the "auto" path makes.
"""

import numpy as np
import pytest
import torch

import teamtomo_amd as tt
from conftest import rel_l2
from oracle import c_oracle as C
from oracle import fourier_slice as O
from teamtomo_amd.synthetic import blob_volume, uniform_rotations, volume_rfft

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture
def binned():
    tt.set_scatter_path("binned")
    yield
    tt.set_scatter_path("auto")


def _tilt_series(tilts_deg, axis_deg):
    """Rotations (xyz) of a single-axis tilt series whose axis lies in the
    image plane at axis_deg from y."""
    out = []
    for t in np.deg2rad(tilts_deg):
        c, s = np.cos(t), np.sin(t)
        tilt = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
        a = np.deg2rad(axis_deg)
        ca, sa = np.cos(a), np.sin(a)
        inplane = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
        out.append(inplane @ tilt @ inplane.T)
    return torch.from_numpy(np.stack(out)).float()


def _normal(shape, seed, device):
    g = torch.Generator().manual_seed(seed)
    return torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g)).to(device)


def _check(S, n, R, fmax, device, label):
    Ri = O.prepare_rotations(R.cpu().numpy(), False, None)
    V, W = tt.insert_central_slices_rfft_3d(S, (n,) * 3, R, fftfreq_max=fmax)
    rV, rW = C.insert(S.cpu().numpy(), (n,) * 3, Ri, fmax)
    e_v, e_w = rel_l2(V.cpu().numpy(), rV), rel_l2(W.cpu().numpy(), rW)
    vol = volume_rfft(blob_volume(n, 30, seed=7, device=device)).requires_grad_(True)
    tt.extract_central_slices_rfft_3d(vol, (n,) * 3, R, fftfreq_max=fmax).backward(S)
    Rz = O.prepare_rotations(R.cpu().numpy(), False, 1e-8)
    ref = C.extract_adjoint(S.cpu().numpy(), tuple(vol.shape), (n,) * 3, Rz, fmax)
    e_a = rel_l2(vol.grad.cpu().numpy(), ref)
    print(f"{label}: insertion volume {e_v:.2e}, weights {e_w:.2e}, adjoint {e_a:.2e}")
    assert e_v < TOL and e_w < TOL and e_a < TOL
    return V, W, vol.grad


@pytest.mark.parametrize("n,fmax", [(96, None), (97, 0.35)])
def test_binned_tilt_series_vs_c_oracle(device, binned, n, fmax):
    R = _tilt_series(np.arange(-60, 61, 2), 7.0).to(device)
    S = _normal((R.shape[0], n, n // 2 + 1), 1, device)
    V, W, gV = _check(S, n, R, fmax, device, f"61 tilts, {n}^3, fftfreq_max={fmax}")
    # bitwise reproducible
    V2, W2 = tt.insert_central_slices_rfft_3d(S, (n,) * 3, R, fftfreq_max=fmax)
    assert torch.equal(V, V2) and torch.equal(W, W2)


def test_binned_exact_geometry_rotations(device, binned):
    """Identity, 90-degree turns and a 180-degree flip: every pixel lands on
    voxels (fractions 0), many on brick faces; plus 45-degree tilts whose
    pixels straddle them."""
    from scipy.spatial.transform import Rotation

    n = 48
    eul = [(0, 0, 0), (90, 0, 0), (0, 90, 0), (0, 0, 90), (180, 0, 0), (90, 90, 0), (0, 45, 0), (45, 0, 45)]
    R = torch.from_numpy(Rotation.from_euler("zyz", eul, degrees=True).as_matrix()).float().to(device)
    S = _normal((R.shape[0], n, n // 2 + 1), 2, device)
    _check(S, n, R, None, device, "exact rotations, 48^3")


def test_binned_many_rotations_vs_c_oracle(device, binned):
    """1500 rotations at 40^3: bins of thousands of records, whose 64-bit sums
    reserve ceil(log2(count)) bits."""
    n = 40
    R = uniform_rotations(1500, 3, device=device)
    S = _normal((1500, n, n // 2 + 1), 4, device)
    _check(S, n, R, None, device, "1500 rotations, 40^3")


def test_binned_multichannel_insert_vs_oracle(device, binned):
    n, c = 40, 3
    R = _tilt_series(np.arange(-60, 61, 6), -20.0).to(device)
    S = _normal((R.shape[0], c, n, n // 2 + 1), 5, device)
    V, W = tt.insert_central_slices_rfft_3d_multichannel(S, (n,) * 3, R)
    rV, rW = O.insert_central_slices_multichannel(S.cpu().numpy(), (n,) * 3,
                                                  O.prepare_rotations(R.cpu().numpy(), False, None))
    e_v, e_w = rel_l2(V.cpu().numpy(), rV), rel_l2(W.cpu().numpy(), rW)
    print(f"multichannel binned insertion: volume {e_v:.2e}, weights {e_w:.2e}")
    assert V.shape == (c, n, n, n // 2 + 1) and e_v < TOL and e_w < TOL


def test_binned_empty_calls(device, binned):
    n = 32
    R = uniform_rotations(0, 6, device=device)
    S = torch.zeros((0, n, n // 2 + 1), dtype=torch.complex64, device=device)
    V, W = tt.insert_central_slices_rfft_3d(S, (n,) * 3, R)
    assert not V.abs().any() and not W.any()
    # a disc too small to hold any pixel but DC
    R = uniform_rotations(5, 7, device=device)
    S = _normal((5, n, n // 2 + 1), 8, device)
    _check(S, n, R, 0.001, device, "fftfreq_max 0.001")


def test_auto_path_choice(device):
    """"auto" bins a 61-tilt series of 512^2 into 512^3 (196 pixels per brick
    of 2048 voxels) and keeps the brick engine for 4096 slices of 64^2 into
    64^3 (70000 per brick): both
    results match the binned and brick results forced by name, bit for bit."""
    n = 64
    R = uniform_rotations(4096, 9, device=device)
    S = _normal((4096, n, n // 2 + 1), 10, device)
    outs = {}
    for path in ("auto", "brick"):
        tt.set_scatter_path(path)
        try:
            outs[path] = tt.insert_central_slices_rfft_3d(S, (n,) * 3, R)
        finally:
            tt.set_scatter_path("auto")
    assert torch.equal(outs["auto"][0], outs["brick"][0]) and torch.equal(outs["auto"][1], outs["brick"][1])
    n = 512
    Rt = _tilt_series(np.arange(-60, 61, 2), 3.0).to(device)
    St = _normal((Rt.shape[0], n, n // 2 + 1), 11, device)
    outs = {}
    for path in ("auto", "binned"):
        tt.set_scatter_path(path)
        try:
            outs[path] = tt.insert_central_slices_rfft_3d(St, (n,) * 3, Rt)
        finally:
            tt.set_scatter_path("auto")
    assert torch.equal(outs["auto"][0], outs["binned"][0]) and torch.equal(outs["auto"][1], outs["binned"][1])


def test_binned_768cubed_tilt_series_vs_c_oracle(device):
    """61 tilts of 768^2 into 768^3 on the default path (the binned scatter:
    a 3.5 GB record buffer, 140 pixels per brick), against the FP64 C oracle."""
    n = 768
    R = _tilt_series(np.arange(-60, 61, 2), -4.0).to(device)
    S = _normal((R.shape[0], n, n // 2 + 1), 12, device)
    V, W = tt.insert_central_slices_rfft_3d(S, (n,) * 3, R)
    tt.set_scatter_path("binned")
    try:
        Vb, Wb = tt.insert_central_slices_rfft_3d(S, (n,) * 3, R)
    finally:
        tt.set_scatter_path("auto")
    assert torch.equal(V, Vb) and torch.equal(W, Wb)  # the default path is the binned scatter
    del Vb, Wb
    rV, rW = C.insert(S.cpu().numpy(), (n,) * 3, O.prepare_rotations(R.cpu().numpy(), False, None))
    e_v, e_w = rel_l2(V.cpu().numpy(), rV), rel_l2(W.cpu().numpy(), rW)
    print(f"61 tilts into 768^3: insertion volume {e_v:.2e}, weights {e_w:.2e}")
    assert e_v < TOL and e_w < TOL
