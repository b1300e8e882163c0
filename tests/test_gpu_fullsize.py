"""Full-size checks of the production configuration: a (256, 256, 129)
complex64 rfft of a 200-blob 256^3 volume and 4096 uniform-SO(3) rotations.

* Insertion of 4096 N(0,1) complex64 slices
  (slice_insertion/_insert_central_slices_rfft_3d.py:9-106): the first 256
  slices against the FP64 C oracle; the whole run through the x = 0 Friedel
  symmetry the mirror (:350-404) gives it, and a bitwise rerun in
  deterministic mode.
* The extraction adjoint (the volume gradient of
  slice_extraction/_extract_central_slices_rfft_3d.py:84-222) of 4096 N(0,1)
  gradients: the first 256 rotations against the FP64 C oracle, the whole run
  through the adjoint identity <A X, g> = <X, A^T g> in FP64 sums for an
  N(0,1) volume X (the blob volume's DC, 1e5 times its mean, would make the
  sum cancel to 5e-5 of its terms and hide nothing but rounding), and a
  bitwise rerun in deterministic mode.
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

N, N_ROT, N_ORACLE = 256, 4096, 256


def _normal_slices(seed, device):
    g = torch.Generator(device=device).manual_seed(seed)
    shape = (N_ROT, N, N // 2 + 1)
    return torch.complex(torch.randn(shape, generator=g, device=device),
                         torch.randn(shape, generator=g, device=device)).contiguous()


def _dot64(a, b):
    """Re sum conj(a) b in FP64."""
    return float(torch.sum(torch.conj(a.to(torch.complex128)) * b.to(torch.complex128)).real)


def test_config2_insertion_4096_slices_256cubed(device):
    R = uniform_rotations(N_ROT, 0, device=device)
    S = _normal_slices(3, device)
    print(f"insertion: {N_ROT} N(0,1) slices into ({N}, {N}, {N // 2 + 1})", flush=True)
    V, W = tt.insert_central_slices_rfft_3d(S, (N, N, N), R)
    torch.cuda.synchronize()

    # the first 256 slices against the FP64 oracle
    Vs, Ws = tt.insert_central_slices_rfft_3d(S[:N_ORACLE], (N, N, N), R[:N_ORACLE])
    Ri = O.prepare_rotations(R[:N_ORACLE].cpu().numpy(), False, None)
    rV, rW = C.insert(S[:N_ORACLE].cpu().numpy(), (N, N, N), Ri)
    e_v, e_w = rel_l2(Vs.cpu().numpy(), rV), rel_l2(Ws.cpu().numpy(), rW)
    print(f"  first {N_ORACLE} slices vs FP64 oracle: volume rel_l2 {e_v:.2e}, weights {e_w:.2e}", flush=True)
    assert e_v < 1e-5 and e_w < 1e-5

    # x = 0 plane: V[z, y, 0] = P[z, y] + conj(P[-z, -y]) for every (z, y) but
    # DC, so mirrored pairs are exact conjugates and self-mirrored voxels real
    P, Pw = V[:, :, 0].cpu(), W[:, :, 0].cpu()
    z = torch.arange(N)
    mz = (N - z) % N
    M, Mw = P[mz][:, mz], Pw[mz][:, mz]
    keep = torch.ones(N, N, dtype=torch.bool)
    keep[N // 2, N // 2] = False
    assert torch.equal(P[keep], torch.conj(M).resolve_conj()[keep]) and torch.equal(Pw[keep], Mw[keep])
    for (a, b) in ((0, 0), (0, N // 2), (N // 2, 0)):
        assert float(P[a, b].imag) == 0.0
    assert float(W.sum()) > 0 and bool(torch.isfinite(V).all())

    # deterministic mode: bitwise equal reruns
    tt.set_deterministic(True)
    try:
        a = tt.insert_central_slices_rfft_3d(S, (N, N, N), R)
        b = tt.insert_central_slices_rfft_3d(S, (N, N, N), R)
    finally:
        tt.set_deterministic(False)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    e_det = rel_l2(a[0].cpu().numpy(), V.cpu().numpy())
    print(f"  deterministic vs default insertion: rel_l2 {e_det:.2e}", flush=True)
    assert e_det < 1e-5


def test_headline_adjoint_4096_rotations_256cubed(device):
    vol = volume_rfft(blob_volume(N, 200, seed=2, device=device)).contiguous()
    R = uniform_rotations(N_ROT, 0, device=device)
    gS = _normal_slices(3, device)
    print(f"extraction + adjoint: ({N}, {N}, {N // 2 + 1}) x {N_ROT} rotations", flush=True)
    V = vol.clone().requires_grad_(True)
    S = tt.extract_central_slices_rfft_3d(V, (N, N, N), R)
    S.backward(gS)
    gV = V.grad.detach()
    torch.cuda.synchronize()

    # adjoint identity over the whole run (FP64 sums of the GPU results)
    g = torch.Generator(device=device).manual_seed(4)
    X = torch.complex(torch.randn(vol.shape, generator=g, device=device),
                      torch.randn(vol.shape, generator=g, device=device)).contiguous()
    lhs = _dot64(tt.extract_central_slices_rfft_3d(X, (N, N, N), R), gS)
    rhs = _dot64(X, gV)
    print(f"  <A V, g> = {lhs:.9e}, <V, A^T g> = {rhs:.9e}, rel diff {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)

    # the first 256 rotations against the FP64 oracle
    V2 = vol.clone().requires_grad_(True)
    tt.extract_central_slices_rfft_3d(V2, (N, N, N), R[:N_ORACLE]).backward(gS[:N_ORACLE])
    Rz = O.prepare_rotations(R[:N_ORACLE].cpu().numpy(), False, 1e-8)
    ref = C.extract_adjoint(gS[:N_ORACLE].cpu().numpy(), vol.shape, (N, N, N), Rz)
    e = rel_l2(V2.grad.cpu().numpy(), ref)
    ref_s = C.extract(vol.cpu().numpy(), (N, N, N), Rz)
    e_s = rel_l2(S[:N_ORACLE].detach().cpu().numpy(), ref_s)
    print(f"  first {N_ORACLE} rotations vs FP64 oracle: slices rel_l2 {e_s:.2e}, volume gradient {e:.2e}",
          flush=True)
    assert e_s < 1e-5 and e < 1e-5

    # deterministic mode: bitwise equal reruns, close to the default result
    tt.set_deterministic(True)
    try:
        grads = []
        for _ in range(2):
            Vd = vol.clone().requires_grad_(True)
            tt.extract_central_slices_rfft_3d(Vd, (N, N, N), R).backward(gS)
            grads.append(Vd.grad)
    finally:
        tt.set_deterministic(False)
    assert torch.equal(grads[0], grads[1])
    e_det = rel_l2(grads[0].cpu().numpy(), gV.cpu().numpy())
    print(f"  deterministic vs default adjoint: rel_l2 {e_det:.2e}", flush=True)
    assert e_det < 1e-5
