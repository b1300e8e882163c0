"""Precision of the scatters in every frequency shell.

A global relative L2 error is dominated by the few large values near DC, so it
can hide a loss of precision at high frequency, where cryo-EM consumers judge
a reconstruction (per-shell FSC).  These tests insert, and apply the
extraction adjoint to, slices of a steep spectrum (the rfft of a 64^3 blob
volume, whose DC is about 1e4 times its Nyquist amplitude) and compare the
result shell by shell with the FP64 oracle on the same complex64 inputs, in
the default (FP32 atomics) and the deterministic (fixed-point) modes.
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


@pytest.mark.parametrize("deterministic", [False, True])
def test_insertion_precision_per_shell(device, deterministic):
    vol, R, S = _steep_slices(device)
    dc, top = np.abs(S).max(), np.abs(S[:, 0, -1]).max()
    print(f"slice spectrum: max |S| {dc:.3e}, largest Nyquist-column |S| {top:.3e} (ratio {dc / top:.1e})")
    tt.set_deterministic(deterministic)
    try:
        V, W = tt.insert_central_slices_rfft_3d(torch.from_numpy(S).to(device), (N,) * 3, R.to(device))
    finally:
        tt.set_deterministic(False)
    rV, rW = O.insert_central_slices(S.astype(np.complex128), (N,) * 3, O.prepare_rotations(R.numpy(), False, None))
    sh = _shells(rV.shape)
    ev = _per_shell(V.cpu().numpy(), rV, sh, N // 2)
    ew = _per_shell(W.cpu().numpy(), rW, sh, N // 2)
    print(f"deterministic={deterministic}: per-shell rel error, volume max {ev.max():.2e} "
          f"(shells 12-31 max {ev[11:].max():.2e}), weights max {ew.max():.2e}")
    assert ev.max() <= SHELL_TOL and ew.max() <= SHELL_TOL


@pytest.mark.parametrize("deterministic", [False, True])
def test_adjoint_precision_per_shell(device, deterministic):
    vol, R, S = _steep_slices(device)
    V = vol.to(device).requires_grad_(True)
    tt.set_deterministic(deterministic)
    try:
        tt.extract_central_slices_rfft_3d(V, (N,) * 3, R.to(device)).backward(torch.from_numpy(S).to(device))
    finally:
        tt.set_deterministic(False)
    Rz = O.prepare_rotations(R.numpy(), False, 1e-8)
    ref = O.extract_adjoint(S.astype(np.complex128), vol.shape, (N,) * 3, Rz)
    e = _per_shell(V.grad.cpu().numpy(), ref, _shells(ref.shape), N // 2)
    print(f"deterministic={deterministic}: adjoint per-shell rel error max {e.max():.2e} "
          f"(shells 12-31 max {e[11:].max():.2e})")
    assert e.max() <= SHELL_TOL
