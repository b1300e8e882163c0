"""Autograd of sample_image_3d and affine_transform_image_3d on the GPU.

The reference differentiates both through F.grid_sample
(torch-image-interpolation/.../image_interpolation_3d.py:64-89,
torch-transform-image/.../transforms_3d.py:22-40); here the sampler's
backward is the tt_sample3d_bwd_* kernel.  Checked against the golden
torch-autograd gradients, the FP64 oracle (sample_image_3d_grad) at a larger
size, FP64 finite differences (torch.autograd.gradcheck on float64 inputs)
and, for the warp, the golden image and matrix gradients.
"""

import numpy as np
import pytest
import torch

import teamtomo_amd as tt
from conftest import golden, rel_l2
from oracle import fourier_slice as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("key", ["real", "complex", "multi"])
@pytest.mark.parametrize("mode", ["trilinear", "nearest"])
def test_sample_grad_vs_golden(device, key, mode):
    G = golden("sample_autograd")
    img = torch.from_numpy(G[f"img_{key}"]).to(device).requires_grad_(True)
    crd = torch.from_numpy(G["coords"]).to(device).requires_grad_(True)
    tgt = torch.from_numpy(G[f"{key}_{mode}_target"]).to(device)
    s = tt.sample_image_3d(img, crd, interpolation=mode)
    loss = torch.real(torch.sum(torch.conj(tgt) * s)) if s.is_complex() else torch.sum(tgt * s)
    loss.backward()
    e_s = rel_l2(s.detach().cpu().numpy(), G[f"{key}_{mode}_samples"])
    e_i = rel_l2(img.grad.cpu().numpy(), G[f"{key}_{mode}_grad_image"])
    e_c = rel_l2(crd.grad.cpu().numpy(), G[f"{key}_{mode}_grad_coords"])
    print(f"{key}/{mode}: samples {e_s:.2e}, grad image {e_i:.2e}, grad coords {e_c:.2e}")
    assert e_s < 1e-6 and e_i < 1e-6 and e_c < 1e-5


@pytest.mark.parametrize("dtype", [torch.float32, torch.complex64, torch.float64])
def test_sample_grad_vs_oracle(device, dtype):
    """2 channels of a 24 x 20 x 28 image, 20000 coordinates (some outside,
    some on integer planes)."""
    rng = np.random.default_rng(11)
    shape = (24, 20, 28)
    real_t = torch.float64 if dtype == torch.float64 else torch.float32
    img = torch.randn((2,) + shape, dtype=real_t, generator=torch.Generator().manual_seed(12))
    if dtype.is_complex:
        img = torch.complex(img, torch.randn(img.shape, generator=torch.Generator().manual_seed(13)))
    crd = rng.uniform(-1.0, 1.0, (20000, 3)) + rng.uniform(0, 1, (20000, 3)) * (np.array(shape) - 1)
    crd[:500] = np.round(crd[:500])  # integer coordinates: one-sided derivatives
    crd_t = torch.from_numpy(crd).to(real_t)
    g = torch.randn((20000, 2), dtype=real_t, generator=torch.Generator().manual_seed(14))
    if dtype.is_complex:
        g = torch.complex(g, torch.randn(g.shape, generator=torch.Generator().manual_seed(15)))
    I = img.to(device).requires_grad_(True)
    c = crd_t.to(device).requires_grad_(True)
    s = tt.sample_image_3d(I, c)
    s.backward(g.to(device))
    prec = np.float64 if dtype == torch.float64 else np.float32
    rI, rc = O.sample_image_3d_grad(img.numpy(), crd_t.numpy(), g.numpy(), precision=prec)
    e_i, e_c = rel_l2(I.grad.cpu().numpy(), rI), rel_l2(c.grad.cpu().numpy(), rc)
    print(f"{dtype}: grad image {e_i:.2e}, grad coords {e_c:.2e} vs FP64 oracle")
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert e_i < tol and e_c < tol


def test_sample_gradcheck_float64(device):
    """torch's own finite-difference check on float64 inputs away from cell faces."""
    g = torch.Generator().manual_seed(16)
    img = torch.randn(2, 5, 6, 7, dtype=torch.float64, generator=g).to(device).requires_grad_(True)
    base = torch.randint(0, 4, (30, 3), generator=g).to(torch.float64)
    crd = (base + 0.1 + 0.8 * torch.rand(30, 3, dtype=torch.float64, generator=g)).to(device).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda a, b: tt.sample_image_3d(a, b), (img, crd), eps=1e-6, atol=1e-8)


def test_affine_3d_grad_vs_golden(device):
    G = golden("sample_autograd")
    img = torch.from_numpy(G["affine_image"]).to(device).requires_grad_(True)
    M = torch.from_numpy(G["affine_matrices"]).to(device).requires_grad_(True)
    tgt = torch.from_numpy(G["affine_target"]).to(device)
    out = tt.affine_transform_image_3d(img, M, "trilinear")
    torch.sum(tgt * out).backward()
    e_o = rel_l2(out.detach().cpu().numpy(), G["affine_warped"])
    e_i = rel_l2(img.grad.cpu().numpy(), G["affine_grad_image"])
    e_m = rel_l2(M.grad.cpu().numpy(), G["affine_grad_matrices"])
    print(f"affine 3-D: warped {e_o:.2e}, grad image {e_i:.2e}, grad matrices {e_m:.2e}")
    assert e_o < 1e-5 and e_i < 1e-5 and e_m < 1e-4
    # the image-only gradient path gives the same image gradient
    img2 = img.detach().clone().requires_grad_(True)
    torch.sum(tgt * tt.affine_transform_image_3d(img2, M.detach(), "trilinear")).backward()
    assert rel_l2(img2.grad.cpu().numpy(), img.grad.cpu().numpy()) < 1e-6
