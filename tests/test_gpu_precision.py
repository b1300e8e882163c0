"""Precision of the scatters in every frequency shell.

A global relative L2 error is dominated by the few large values near DC, so it
can hide a loss of precision at high frequency, where cryo-EM consumers judge
a reconstruction (per-shell FSC).  These tests insert, and apply the
extraction adjoint to, slices of a steep spectrum (the rfft of a 64^3 blob
volume, whose DC is about 1e4 times its Nyquist amplitude) and compare the
result shell by shell with the FP64 oracle on the same complex64 inputs, for
both fixed-point scatters: the brick engine (per-brick scales from sampled
band maxima, 32- or 64-bit contributions) and the binned scatter (per-brick
scales from the exact maxima of its records, 64-bit contributions).
"""

import numpy as np
import pytest
import torch

import teamtomo_amd as tt
from oracle import fourier_slice as O
from teamtomo_amd.synthetic import blob_volume, uniform_rotations, volume_rfft

pytestmark = pytest.mark.gpu

N, N_ROT = 64, 200
SHELL_TOL = 1e-5


def _shells(shape):
    D, H, Wr = shape
    z = np.arange(D)[:, None, None] - D // 2
    y = np.arange(H)[None, :, None] - H // 2
    x = np.arange(Wr)[None, None, :]
    return np.floor(np.sqrt(z * z + y * y + x * x)).astype(int)


def _per_shell(got, ref, shells, top):
    out = []
    for k in range(1, top):
        m = shells == k
        den = np.linalg.norm(ref[m])
        out.append(np.linalg.norm((got - ref)[m]) / den if den > 0 else 0.0)
    return np.array(out)


def _steep_slices(device):
    vol = volume_rfft(blob_volume(N, 200, seed=40))
    R = uniform_rotations(N_ROT, 41)
    Rz = O.prepare_rotations(R.numpy(), False, 1e-8)
    S = O.extract_central_slices(vol.numpy().astype(np.complex128), (N,) * 3, Rz).astype(np.complex64)
    return vol, R, S


@pytest.mark.parametrize("path", ["brick", "binned"])
def test_insertion_precision_per_shell(device, path):
    vol, R, S = _steep_slices(device)
    dc, top = np.abs(S).max(), np.abs(S[:, 0, -1]).max()
    print(f"slice spectrum: max |S| {dc:.3e}, largest Nyquist-column |S| {top:.3e} (ratio {dc / top:.1e})")
    tt.set_scatter_path(path)
    try:
        V, W = tt.insert_central_slices_rfft_3d(torch.from_numpy(S).to(device), (N,) * 3, R.to(device))
    finally:
        tt.set_scatter_path("auto")
    rV, rW = O.insert_central_slices(S.astype(np.complex128), (N,) * 3, O.prepare_rotations(R.numpy(), False, None))
    sh = _shells(rV.shape)
    ev = _per_shell(V.cpu().numpy(), rV, sh, N // 2)
    ew = _per_shell(W.cpu().numpy(), rW, sh, N // 2)
    print(f"{path}: per-shell rel error, volume max {ev.max():.2e} "
          f"(shells 12-31 max {ev[11:].max():.2e}), weights max {ew.max():.2e}")
    assert ev.max() <= SHELL_TOL and ew.max() <= SHELL_TOL


@pytest.mark.parametrize("path", ["brick", "binned"])
def test_adjoint_precision_per_shell(device, path):
    vol, R, S = _steep_slices(device)
    V = vol.to(device).requires_grad_(True)
    tt.set_scatter_path(path)
    try:
        tt.extract_central_slices_rfft_3d(V, (N,) * 3, R.to(device)).backward(torch.from_numpy(S).to(device))
    finally:
        tt.set_scatter_path("auto")
    Rz = O.prepare_rotations(R.numpy(), False, 1e-8)
    ref = O.extract_adjoint(S.astype(np.complex128), vol.shape, (N,) * 3, Rz)
    e = _per_shell(V.grad.cpu().numpy(), ref, _shells(ref.shape), N // 2)
    print(f"{path}: adjoint per-shell rel error max {e.max():.2e} "
          f"(shells 12-31 max {e[11:].max():.2e})")
    assert e.max() <= SHELL_TOL


@pytest.mark.parametrize("n_rot", [200, 1100])
def test_guarded_redo_after_an_unsampled_spike(device, n_rot):
    """A value 1e6 times the rest sits in a slice the sampled band maxima skip
    (every n_rot // 64-th slice is sampled): the brick engine's first pass
    sees it exceed its brick's limit and the call is redone with the exact
    maximum.  Insertion and adjoint must still match the FP64 oracle, with few
    rotations (the brick-crossing pre-pass) and many."""
    from oracle import c_oracle as C

    tt.set_scatter_path("brick")
    try:
        _spike_case(device, n_rot, C)
    finally:
        tt.set_scatter_path("auto")


def test_binned_scatter_with_a_spike(device):
    """The binned scatter takes each brick's exact maximum: a 1e6 spike only
    coarsens the bricks its pixel reaches."""
    from oracle import c_oracle as C

    tt.set_scatter_path("binned")
    try:
        e_far, e_afar = _spike_case(device, 200, C)
    finally:
        tt.set_scatter_path("auto")
    assert e_far < 1e-5 and e_afar < 1e-5


def _spike_case(device, n_rot, C):
    n = 48
    R = uniform_rotations(n_rot, 50, device=device)
    g = torch.Generator(device=device).manual_seed(51)
    shape = (n_rot, n, n // 2 + 1)
    S = torch.complex(torch.randn(shape, generator=g, device=device), torch.randn(shape, generator=g, device=device))
    S[1, n // 2 + 3, 5] = 1e6 + 2e5j  # slice 1 is never sampled
    V, W = tt.insert_central_slices_rfft_3d(S, (n,) * 3, R)
    V2, W2 = tt.insert_central_slices_rfft_3d(S, (n,) * 3, R)
    assert torch.equal(V, V2) and torch.equal(W, W2)
    Ri = O.prepare_rotations(R.cpu().numpy(), False, None)
    rV, rW = C.insert(S.cpu().numpy(), (n,) * 3, Ri)
    e_v = float(np.linalg.norm(V.cpu().numpy() - rV) / np.linalg.norm(rV))
    far = np.abs(rV) < 1e3  # the voxels away from the spike, on their own
    e_far = float(np.linalg.norm((V.cpu().numpy() - rV)[far]) / np.linalg.norm(rV[far]))
    vol = volume_rfft(blob_volume(n, 20, seed=52)).to(device).requires_grad_(True)
    tt.extract_central_slices_rfft_3d(vol, (n,) * 3, R).backward(S)
    Rz = O.prepare_rotations(R.cpu().numpy(), False, 1e-8)
    ref = C.extract_adjoint(S.cpu().numpy(), vol.shape, (n,) * 3, Rz)
    e_a = float(np.linalg.norm(vol.grad.cpu().numpy() - ref) / np.linalg.norm(ref))
    far_a = np.abs(ref) < 1e3
    e_afar = float(np.linalg.norm((vol.grad.cpu().numpy() - ref)[far_a]) / np.linalg.norm(ref[far_a]))
    print(f"{n_rot} rotations with a spike: insert rel_l2 {e_v:.2e} (away from it {e_far:.2e}), "
          f"adjoint {e_a:.2e} (away {e_afar:.2e})")
    assert e_v < 1e-5 and e_a < 1e-5
    # the brick engine's redo scale is the global maximum: 2^-30 of 1.2e6,
    # about 1e-3, per contribution
    assert e_far < 1e-2 and e_afar < 1e-2
    return e_far, e_afar


@pytest.mark.parametrize("signs", ["positive", "negative", "mixed"])
def test_split_accumulator_carries(device, signs):
    """1024 copies of one rotation: every voxel near its plane sums about a
    thousand contributions.  The sampled slices (every 16th) hold values of
    magnitude about 1 and the others values about 200 times larger (under the
    2^8 headroom above the sampled band maxima, so there is no redo), so a
    32-bit contribution is about 2^28.6 and the low words of the brick
    engine's split accumulators wrap many times in every unit: carries for
    positive values, borrows for negative ones, both for mixed signs.  A
    wrong carry would move a voxel by 2^32 quanta (about 2048).  Insertion and
    adjoint must match the FP64 oracle."""
    from conftest import rel_l2
    from oracle import c_oracle as C

    n, k = 64, 1024
    R = uniform_rotations(1, 60).expand(k, 3, 3).contiguous()
    g = torch.Generator().manual_seed(61)
    sign = {"positive": torch.ones(k), "negative": -torch.ones(k),
            "mixed": torch.randint(0, 2, (k,), generator=g).float() * 2.0 - 1.0}[signs]
    mag = torch.full((k,), 200.0)
    mag[:: k // 64] = 1.0  # the slices the band maxima sample
    jitter = 1.0 + 0.1 * torch.rand((k, n, n // 2 + 1), generator=g)
    S = torch.complex(jitter, 0.5 * jitter) * (sign * mag).view(k, 1, 1)
    Sd, Rd = S.to(device), R.to(device)
    tt.set_scatter_path("brick")
    try:
        V, W = tt.insert_central_slices_rfft_3d(Sd, (n,) * 3, Rd)
        vol = torch.zeros((n, n, n // 2 + 1), dtype=torch.complex64, device=device, requires_grad=True)
        tt.extract_central_slices_rfft_3d(vol, (n,) * 3, Rd).backward(Sd)
    finally:
        tt.set_scatter_path("auto")
    rV, rW = C.insert(S.numpy(), (n,) * 3, O.prepare_rotations(R.numpy(), False, None))
    ref = C.extract_adjoint(S.numpy(), tuple(vol.shape), (n,) * 3, O.prepare_rotations(R.numpy(), False, 1e-8))
    e_v, e_w = rel_l2(V.cpu().numpy(), rV), rel_l2(W.cpu().numpy(), rW)
    e_a = rel_l2(vol.grad.cpu().numpy(), ref)
    print(f"{signs}: insertion volume {e_v:.2e}, weights {e_w:.2e}, adjoint {e_a:.2e}; "
          f"largest |V| {np.abs(rV).max():.3e}")
    assert e_v < 1e-6 and e_w < 1e-6 and e_a < 1e-6
