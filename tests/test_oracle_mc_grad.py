"""Pinning of the oracle's multichannel operators and sampler gradients.

* extract / insert / project multichannel against the golden fixtures written
  from the reference (``multichannel_16``) and against the single-channel
  oracle applied per channel;
* sample_image_3d_grad against the golden torch-autograd gradients
  (``sample_autograd``), the adjoint identity of its image gradient and FP64
  central differences of its coordinate gradient;
* affine_transform_image_3d_grad against the golden gradients and FP64
  central differences of the matrices.
"""

import numpy as np
import pytest

from conftest import golden, rel_l2
from oracle import fourier_slice as O
from oracle import transform_image as T


def test_multichannel_oracle_vs_golden():
    G = golden("multichannel_16")
    V, R = G["volume_rfft"], G["rotations"]
    Rz = O.prepare_rotations(R, False, 1e-8)
    S = O.extract_central_slices_multichannel(V, (16, 16, 16), Rz)
    assert S.shape == G["slices"].shape
    assert rel_l2(S, G["slices"]) < 1e-5
    assert rel_l2(O.extract_central_slices_multichannel(V, (16, 16, 16), Rz, 0.3), G["slices_fmax"]) < 1e-5
    for k in range(V.shape[0]):  # the same as the single-channel oracle per channel
        assert np.array_equal(S[:, k], O.extract_central_slices(V[k], (16, 16, 16), Rz))
    Ri = O.prepare_rotations(R, False, None)
    Vi, Wi = O.insert_central_slices_multichannel(G["ins_slices"], (16, 16, 16), Ri)
    assert rel_l2(Vi, G["ins_volume"]) < 1e-5 and rel_l2(Wi, G["ins_weights"]) < 1e-6
    Vf, Wf = O.insert_central_slices_multichannel(G["ins_slices"], (16, 16, 16), Ri, 0.3)
    assert rel_l2(Vf, G["ins_volume_fmax"]) < 1e-5 and rel_l2(Wf, G["ins_weights_fmax"]) < 1e-6
    P = O.project_3d_to_2d_multichannel(G["proj_volume"], G["proj_rotations"])
    assert P.shape == G["projections"].shape and rel_l2(P, G["projections"]) < 1e-5


@pytest.mark.parametrize("key", ["real", "complex", "multi"])
@pytest.mark.parametrize("mode", ["trilinear", "nearest"])
def test_sampler_grad_vs_golden(key, mode):
    G = golden("sample_autograd")
    img, crd, tgt = G[f"img_{key}"], G["coords"], G[f"{key}_{mode}_target"]
    gI, gc = O.sample_image_3d_grad(img, crd, tgt, mode)
    assert rel_l2(gI, G[f"{key}_{mode}_grad_image"]) < 1e-5
    if mode == "nearest":
        assert not np.any(gc) and not np.any(G[f"{key}_{mode}_grad_coords"])
    else:
        assert rel_l2(gc, G[f"{key}_{mode}_grad_coords"]) < 1e-5


@pytest.mark.parametrize("key", ["real", "complex", "multi"])
def test_sampler_grad_adjoint_and_finite_differences(key):
    rng = np.random.default_rng(7)
    shape = (9, 7, 8)
    img = rng.standard_normal(((2,) if key == "multi" else ()) + shape)
    if key == "complex":
        img = img + 1j * rng.standard_normal(shape)
    # coordinates away from integers and the border, where the sampler is smooth
    crd = rng.uniform(0.05, 0.95, (200, 3)) + rng.integers(0, np.array(shape) - 1, (200, 3))
    lead = (200, 2) if key == "multi" else (200,)
    g = rng.standard_normal(lead) + (1j * rng.standard_normal(lead) if key == "complex" else 0)
    gI, gc = O.sample_image_3d_grad(img, crd, g, precision=np.float64)
    # image gradient: the exact adjoint of the (linear in the image) sampler
    s = O.sample_image_3d(img, crd)
    assert abs(np.real(np.vdot(g, s)) - np.real(np.vdot(gI, img))) < 1e-10 * np.abs(g).sum()
    # coordinate gradient: central differences of Re <g, s(c)>
    h = 1e-6
    fd = np.zeros_like(crd)
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        fd[:, d] = (np.real(np.sum(np.conj(g) * O.sample_image_3d(img, crd + e), axis=tuple(range(1, g.ndim))))
                    - np.real(np.sum(np.conj(g) * O.sample_image_3d(img, crd - e), axis=tuple(range(1, g.ndim))))) / (2 * h)
    assert rel_l2(gc, fd) < 1e-8
    # outside points and points on the border pass no coordinate gradient
    edge = np.array([[-0.5, 3.2, 3.3], [0.0, 3.2, 3.3], [8.0, 3.2, 3.3], [4.5, 6.0, 2.2]])
    g1 = np.ones((4, 2)) if key == "multi" else np.ones(4)
    _, gce = O.sample_image_3d_grad(img, edge, g1, precision=np.float64)
    assert not np.any(gce[0]) and gce[1, 0] == 0 and gce[2, 0] == 0 and gce[3, 1] == 0


def test_affine_grad_vs_golden_and_finite_differences():
    G = golden("sample_autograd")
    img, M, tgt = G["affine_image"], G["affine_matrices"], G["affine_target"]
    gI, gM = T.affine_transform_image_3d_grad(img, M, tgt, "trilinear")
    assert rel_l2(gI, G["affine_grad_image"]) < 1e-5
    assert rel_l2(gM, G["affine_grad_matrices"]) < 1e-4
    # FP64 central differences of the loss in a few matrix entries (FP64 matrices)
    M64 = M.astype(np.float64)
    _, gM64 = T.affine_transform_image_3d_grad(img, M64, tgt, "trilinear", exact_matrices=True)
    h = 1e-7
    for idx in [(0, 0, 0, 0, 0, 1), (1, 0, 0, 0, 2, 3), (2, 0, 0, 0, 1, 2)]:
        Mp, Mm = M64.copy(), M64.copy()
        Mp[idx] += h
        Mm[idx] -= h

        def loss(Mx):
            coords = _affine_coords(Mx, img.shape[-3:])
            return float(np.sum(tgt * O.sample_image_3d(img, coords)))

        fd = (loss(Mp) - loss(Mm)) / (2 * h)
        assert abs(fd - gM64[idx]) <= 1e-5 * max(1.0, abs(fd)), (idx, fd, gM64[idx])


def _affine_coords(M, shape):
    M = M.copy()
    M[..., :3, :3] = M[..., :3, :3][..., ::-1, ::-1]
    M[..., :3, 3] = M[..., :3, 3][..., ::-1]
    grid = np.moveaxis(np.indices(shape, dtype=np.float64), 0, -1)
    hom = np.concatenate([grid, np.ones(grid.shape[:-1] + (1,))], axis=-1)[..., None]
    return (M @ hom)[..., :3, 0]
