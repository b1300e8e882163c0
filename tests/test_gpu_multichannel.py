"""Multichannel operators (slice_extraction/_extract_central_slices_rfft_3d.py:225-367,
slice_insertion/_insert_central_slices_rfft_3d.py:109-210, project.py:96-177).

* Extraction: the two-channel brick gather against the single-channel
  extraction of each channel (the same enumeration and arithmetic: bitwise
  equal), for even and odd channel counts, with the fftfreq_max disc, on the
  pixel-gather path, and through autograd.
* Extraction, insertion (weights shared by the channels) and projection
  against the golden fixtures written from the reference and against the FP64
  oracle applied per channel.
"""

import time

import numpy as np
import pytest
import torch

import teamtomo_amd as tt
from conftest import golden, rel_l2
from oracle import c_oracle as C_
from oracle import fourier_slice as O
from teamtomo_amd.synthetic import blob_volume, uniform_rotations, volume_rfft

pytestmark = pytest.mark.gpu


def _vols(C, N, device, seed=40):
    return torch.stack([volume_rfft(blob_volume(N, 30, seed=seed + c, device=device)) for c in range(C)]).contiguous()


@pytest.mark.parametrize("C,N,fmax", [(2, 64, None), (3, 64, 0.3), (4, 48, None), (1, 32, None)])
def test_multichannel_equals_per_channel(device, C, N, fmax):
    V = _vols(C, N, device)
    R = uniform_rotations(70, 41, device=device).reshape(7, 10, 3, 3)
    out = tt.extract_central_slices_rfft_3d_multichannel(V, (N, N, N), R, fftfreq_max=fmax)
    assert out.shape == (7, 10, C, N, N // 2 + 1)
    for c in range(C):
        ref = tt.extract_central_slices_rfft_3d(V[c], (N, N, N), R, fftfreq_max=fmax)
        assert torch.equal(out[..., c, :, :], ref), f"channel {c}"


def test_multichannel_pixel_path_and_image_shape(device):
    V = _vols(3, 40, device, seed=50)
    R = uniform_rotations(33, 51, device=device)
    tt.set_gather_path("pixel")
    try:
        out = tt.extract_central_slices_rfft_3d_multichannel(V, (80, 80, 80), R)
        ref = [tt.extract_central_slices_rfft_3d(V[c], (80, 80, 80), R) for c in range(3)]
    finally:
        tt.set_gather_path("auto")
    for c in range(3):
        assert torch.equal(out[:, c], ref[c])
    brick = tt.extract_central_slices_rfft_3d_multichannel(V, (80, 80, 80), R)
    scale = float(out.abs().max())
    assert float((brick - out).abs().max()) <= 1e-6 * scale


def test_multichannel_backward(device):
    V = _vols(3, 24, device, seed=60).requires_grad_(True)
    R = uniform_rotations(9, 61, device=device).to(torch.float64).requires_grad_(True)
    g = torch.Generator(device=device).manual_seed(62)
    T = torch.complex(torch.randn((9, 3, 24, 13), generator=g, device=device),
                      torch.randn((9, 3, 24, 13), generator=g, device=device))
    out = tt.extract_central_slices_rfft_3d_multichannel(V, (24, 24, 24), R)
    torch.real(torch.sum(torch.conj(T) * out)).backward()
    gV, gR = V.grad.clone(), R.grad.clone()
    V2 = V.detach().clone().requires_grad_(True)
    R2 = R.detach().clone().requires_grad_(True)
    loss = sum(torch.real(torch.sum(torch.conj(T[:, c]) * tt.extract_central_slices_rfft_3d(V2[c], (24, 24, 24), R2)))
               for c in range(3))
    loss.backward()
    assert torch.allclose(gV, V2.grad, rtol=0, atol=1e-5 * float(V2.grad.abs().max()))
    # the fused backward sums the channels' rotation gradients in FP64, the
    # per-channel loop in float32 through autograd: equal to float32 rounding of the sum
    assert torch.allclose(gR, R2.grad, rtol=0, atol=1e-6 * float(R2.grad.abs().max()))


def test_multichannel_vs_golden(device):
    G = golden("multichannel_16")
    V = torch.from_numpy(G["volume_rfft"]).to(device)
    R = torch.from_numpy(G["rotations"]).to(device)
    S = tt.extract_central_slices_rfft_3d_multichannel(V, (16, 16, 16), R)
    Sf = tt.extract_central_slices_rfft_3d_multichannel(V, (16, 16, 16), R, fftfreq_max=0.3)
    D = torch.from_numpy(G["ins_slices"]).to(device)
    Vi, Wi = tt.insert_central_slices_rfft_3d_multichannel(D, (16, 16, 16), R)
    Vf, Wf = tt.insert_central_slices_rfft_3d_multichannel(D, (16, 16, 16), R, fftfreq_max=0.3)
    P = tt.project_3d_to_2d_multichannel(torch.from_numpy(G["proj_volume"]).to(device),
                                         torch.from_numpy(G["proj_rotations"]).to(device))
    errs = {
        "extract": rel_l2(S.cpu().numpy(), G["slices"]), "extract fmax": rel_l2(Sf.cpu().numpy(), G["slices_fmax"]),
        "insert": rel_l2(Vi.cpu().numpy(), G["ins_volume"]), "weights": rel_l2(Wi.cpu().numpy(), G["ins_weights"]),
        "insert fmax": rel_l2(Vf.cpu().numpy(), G["ins_volume_fmax"]),
        "weights fmax": rel_l2(Wf.cpu().numpy(), G["ins_weights_fmax"]),
        "project": rel_l2(P.cpu().numpy(), G["projections"]),
    }
    print("multichannel vs golden:", {k: f"{v:.1e}" for k, v in errs.items()})
    assert S.shape == G["slices"].shape and Vi.shape == G["ins_volume"].shape and P.shape == G["projections"].shape
    assert max(errs.values()) < 1e-5


@pytest.mark.parametrize("C,N,fmax", [(3, 48, None), (2, 64, 0.35)])
def test_multichannel_vs_oracle(device, C, N, fmax):
    """Extraction and insertion (shared weights) of C channels against the
    FP64 C oracle per channel, and the insertion's backward against the
    per-channel single-channel backward."""
    V = _vols(C, N, device, seed=80)
    R = uniform_rotations(64, 81, device=device)
    Rz = O.prepare_rotations(R.cpu().numpy(), False, 1e-8)
    Ri = O.prepare_rotations(R.cpu().numpy(), False, None)
    S = tt.extract_central_slices_rfft_3d_multichannel(V, (N, N, N), R, fftfreq_max=fmax)
    Vn = V.cpu().numpy()
    e_x = max(rel_l2(S[:, c].cpu().numpy(), C_.extract(Vn[c], (N, N, N), Rz, fmax)) for c in range(C))
    g = torch.Generator(device=device).manual_seed(82)
    D = torch.complex(torch.randn(S.shape, generator=g, device=device), torch.randn(S.shape, generator=g, device=device))
    Vi, Wi = tt.insert_central_slices_rfft_3d_multichannel(D, (N, N, N), R, fftfreq_max=fmax)
    e_v, e_w = 0.0, 0.0
    for c in range(C):
        rV, rW = C_.insert(D[:, c].cpu().numpy(), (N, N, N), Ri, fmax)
        e_v = max(e_v, rel_l2(Vi[c].cpu().numpy(), rV))
        e_w = max(e_w, rel_l2(Wi.cpu().numpy(), rW))
    print(f"{C} channels {N}^3 x 64, fmax {fmax}: vs FP64 oracle extract {e_x:.2e}, insert {e_v:.2e}, "
          f"weights {e_w:.2e}")
    assert e_x < 1e-5 and e_v < 1e-5 and e_w < 1e-6
    # backward of the multichannel insertion: slices and rotations
    Dg = D.clone().requires_grad_(True)
    Rg = R.to(torch.float64).clone().requires_grad_(True)
    TV = torch.complex(torch.randn(Vi.shape, generator=g, device=device), torch.randn(Vi.shape, generator=g, device=device))
    TW = torch.randn(Wi.shape, generator=g, device=device)
    Vg, Wg = tt.insert_central_slices_rfft_3d_multichannel(Dg, (N, N, N), Rg, fftfreq_max=fmax)
    (torch.real(torch.sum(torch.conj(TV) * Vg)) + torch.sum(TW * Wg)).backward()
    D2 = D.clone().requires_grad_(True)
    R2 = R.to(torch.float64).clone().requires_grad_(True)
    loss = 0.0
    for c in range(C):
        v, w = tt.insert_central_slices_rfft_3d(D2[:, c], (N, N, N), R2, fftfreq_max=fmax)
        loss = loss + torch.real(torch.sum(torch.conj(TV[c]) * v)) + (torch.sum(TW * w) if c == 0 else 0.0)
    loss.backward()
    assert rel_l2(Dg.grad.cpu().numpy(), D2.grad.cpu().numpy()) < 1e-6
    assert rel_l2(Rg.grad.cpu().numpy(), R2.grad.cpu().numpy()) < 1e-5


def test_multichannel_projection_vs_oracle(device):
    vol = torch.stack([blob_volume(20, 8, seed=90 + c) for c in range(3)])
    R = uniform_rotations(12, 93)
    P = tt.project_3d_to_2d_multichannel(vol.to(device), R.to(device))
    ref = O.project_3d_to_2d_multichannel(vol.numpy(), R.numpy())
    e = rel_l2(P.cpu().numpy(), ref)
    print(f"multichannel projection 3 x 20^3 pad 2, 12 poses: rel_l2 vs FP64 oracle {e:.2e}")
    assert P.shape == (12, 3, 20, 20) and e < 1e-5


def test_multichannel_timing(device):
    V = _vols(4, 128, device, seed=70)
    R = uniform_rotations(1024, 71, device=device)

    def fused():
        return tt.extract_central_slices_rfft_3d_multichannel(V, (128, 128, 128), R)

    def loop():
        return [tt.extract_central_slices_rfft_3d(V[c], (128, 128, 128), R) for c in range(4)]

    for f in (fused, loop):
        f()
    torch.cuda.synchronize()
    t = {}
    for name, f in (("fused", fused), ("loop", loop)):
        t0 = time.perf_counter()
        for _ in range(5):
            f()
        torch.cuda.synchronize()
        t[name] = (time.perf_counter() - t0) / 5 * 1e3
    print(f"4-channel 128^3 x 1024 extraction: fused {t['fused']:.2f} ms, per-channel loop {t['loop']:.2f} ms "
          f"({t['loop'] / t['fused']:.2f}x)")
    # two channels per brick launch share the enumeration of the slice pixels
    assert t["fused"] < t["loop"]
