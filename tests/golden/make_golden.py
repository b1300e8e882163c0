"""Write the golden input/output vectors of the reference Fourier-slice path.

Run once, on a machine that has the reference sources, with
``python tests/golden/make_golden.py [REFERENCE_ROOT] [--2d]`` (default
``/root/reference``; ``--2d`` / ``--transform`` / ``--mc-grad`` write only the
fixtures of the 2-D analogues / of the affine warps / of the multichannel
operators and the sampler's autograd).  It imports the reference's pure-Python packages
(torch-fourier-slice, torch-image-interpolation, torch-grid-utils) on the CPU
and stores inputs and outputs as small ``.npz`` fixtures next to this script.
Nothing at test time reads the reference: the tests load the fixtures.

The fixtures pin the composite float32 behaviour that closed forms and
invariants leave open (odd box sizes, an image_shape different from the
volume shape, the fftfreq_max disc, the insertion Friedel mirror, nearest
rounding ties, torch-autograd gradients, the near-zero conjugation tie-break).
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _import_reference(root):
    pkgs = os.path.join(root, "packages", "primitives")
    for name in ("torch-fourier-slice", "torch-image-interpolation", "torch-grid-utils", "torch-affine-utils",
                 "torch-transform-image", "torch-fourier-shift", "torch-subpixel-crop"):
        sys.path.insert(0, os.path.join(pkgs, name, "src"))
    import torch_fourier_slice as tfs  # noqa: E402
    import torch_image_interpolation as tii  # noqa: E402

    return tfs, tii


def uniform_rotations(n, seed):
    """Uniform SO(3) matrices from normalised Gaussian quaternions (float32)."""
    import torch

    g = torch.Generator().manual_seed(seed)
    q = torch.randn(n, 4, generator=g, dtype=torch.float64)
    q = q / q.norm(dim=-1, keepdim=True)
    w, x, y, z = q.unbind(-1)
    R = torch.stack(
        [
            1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
            2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
            2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y),
        ],
        dim=-1,
    ).reshape(n, 3, 3)
    return R.to(torch.float32)


def zyz_matrix(angles_deg):
    """Intrinsic ZYZ Euler angles (degrees) -> xyz rotation matrix, float64.

    Equals scipy's Rotation.from_euler("ZYZ", angles, degrees=True).as_matrix().
    """
    from scipy.spatial.transform import Rotation

    return Rotation.from_euler("ZYZ", angles_deg, degrees=True).as_matrix()


def vol_rfft(v):
    import torch

    f = torch.fft.fftshift(v, dim=(-3, -2, -1))
    f = torch.fft.rfftn(f, dim=(-3, -2, -1))
    return torch.fft.fftshift(f, dim=(-3, -2))


def main(root):
    import torch

    torch.set_num_threads(8)
    tfs, tii = _import_reference(root)
    ext = tfs.extract_central_slices_rfft_3d
    ins = tfs.insert_central_slices_rfft_3d

    def save(name, **arrays):
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
        print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")

    # G1/G2: 64^3 N(0,1) volume, 100 uniform rotations; extract then insert.
    g = torch.Generator().manual_seed(1)
    v = torch.randn(64, 64, 64, generator=g)
    V = vol_rfft(v)
    R = uniform_rotations(100, 0)
    S = ext(V, (64, 64, 64), R.clone())
    Vi, Wi = ins(S, (64, 64, 64), R.clone())
    save("extract_insert_64", volume_rfft=V.numpy(), rotations=R.numpy(), slices=S.numpy(),
         ins_volume=Vi.numpy(), ins_weights=Wi.numpy())

    # G3: odd box 33^3.
    g = torch.Generator().manual_seed(5)
    v = torch.randn(33, 33, 33, generator=g)
    V = vol_rfft(v)
    R = uniform_rotations(20, 6)
    S = ext(V, (33, 33, 33), R.clone())
    Vi, Wi = ins(S, (33, 33, 33), R.clone())
    save("odd_33", volume_rfft=V.numpy(), rotations=R.numpy(), slices=S.numpy(),
         ins_volume=Vi.numpy(), ins_weights=Wi.numpy())

    # G4: image_shape 4x the volume shape.
    g = torch.Generator().manual_seed(7)
    v = torch.randn(16, 16, 16, generator=g)
    V = vol_rfft(v)
    R = uniform_rotations(12, 8)
    S = ext(V, (64, 64, 64), R.clone())
    save("image_shape_4x", volume_rfft=V.numpy(), rotations=R.numpy(), slices=S.numpy())

    # G5: fftfreq_max = 0.3 at 32^3.
    g = torch.Generator().manual_seed(9)
    v = torch.randn(32, 32, 32, generator=g)
    V = vol_rfft(v)
    R = uniform_rotations(16, 10)
    S = ext(V, (32, 32, 32), R.clone(), fftfreq_max=0.3)
    Vi, Wi = ins(S, (32, 32, 32), R.clone(), fftfreq_max=0.3)
    save("fftfreq_max_03", volume_rfft=V.numpy(), rotations=R.numpy(), slices=S.numpy(),
         ins_volume=Vi.numpy(), ins_weights=Wi.numpy())

    # G6: generic sampler, real / complex / multichannel, trilinear / nearest,
    # coordinates partly out of bounds and exact .5 ties.
    g = torch.Generator().manual_seed(11)
    img_r = torch.randn(12, 10, 9, generator=g)
    img_c = torch.complex(torch.randn(12, 10, 9, generator=g), torch.randn(12, 10, 9, generator=g))
    img_m = torch.randn(3, 12, 10, 9, generator=g)
    hi = torch.tensor([12.0, 10.0, 9.0])
    crd = torch.rand(400, 3, generator=g) * (hi + 1.0) - 1.0
    ties = torch.randint(0, 8, (100, 3), generator=g).float() + 0.5
    ints = torch.randint(0, 9, (50, 3), generator=g).float()
    crd = torch.cat([crd, ties, ints]).reshape(11, 50, 3)
    out = {}
    for mode in ("trilinear", "nearest"):
        out[f"real_{mode}"] = tii.sample_image_3d(img_r, crd, interpolation=mode).numpy()
        out[f"complex_{mode}"] = tii.sample_image_3d(img_c, crd, interpolation=mode).numpy()
        out[f"multi_{mode}"] = tii.sample_image_3d(img_m, crd, interpolation=mode).numpy()
    save("sample_image_3d", image_real=img_r.numpy(), image_complex=img_c.numpy(),
         image_multi=img_m.numpy(), coordinates=crd.numpy(), **out)

    # G7: generic splat, trilinear / nearest, two accumulating calls.
    g = torch.Generator().manual_seed(12)
    crd1 = torch.rand(300, 3, generator=g) * (hi + 1.0) - 1.0
    crd1 = torch.cat([crd1, ties[:50]])
    crd2 = torch.rand(200, 3, generator=g) * (hi + 1.0) - 1.0
    val_r1 = torch.randn(350, generator=g)
    val_r2 = torch.randn(200, generator=g)
    val_c1 = torch.complex(torch.randn(350, generator=g), torch.randn(350, generator=g))
    val_c2 = torch.complex(torch.randn(200, generator=g), torch.randn(200, generator=g))
    val_m1 = torch.randn(350, 2, generator=g)
    val_m2 = torch.randn(200, 2, generator=g)
    out = {}
    for mode in ("trilinear", "nearest"):
        im, w = tii.insert_into_image_3d(val_r1, crd1, torch.zeros(12, 10, 9), interpolation=mode)
        im, w = tii.insert_into_image_3d(val_r2, crd2, im, w, interpolation=mode)
        out[f"real_{mode}_image"], out[f"real_{mode}_weights"] = im.numpy(), w.numpy()
        im, w = tii.insert_into_image_3d(val_c1, crd1, torch.zeros(12, 10, 9, dtype=torch.complex64),
                                         interpolation=mode)
        im, w = tii.insert_into_image_3d(val_c2, crd2, im, w, interpolation=mode)
        out[f"complex_{mode}_image"], out[f"complex_{mode}_weights"] = im.numpy(), w.numpy()
        im, w = tii.insert_into_image_3d(val_m1, crd1, torch.zeros(2, 12, 10, 9), interpolation=mode)
        im, w = tii.insert_into_image_3d(val_m2, crd2, im, w, interpolation=mode)
        out[f"multi_{mode}_image"], out[f"multi_{mode}_weights"] = im.numpy(), w.numpy()
    save("insert_into_image_3d", coords1=crd1.numpy(), coords2=crd2.numpy(),
         val_r1=val_r1.numpy(), val_r2=val_r2.numpy(), val_c1=val_c1.numpy(),
         val_c2=val_c2.numpy(), val_m1=val_m1.numpy(), val_m2=val_m2.numpy(), **out)

    # G8: project / backproject of a 24^3 blob volume, pad 2.
    g = torch.Generator().manual_seed(13)
    zz, yy, xx = torch.meshgrid(*[torch.arange(24.0)] * 3, indexing="ij")
    vol = torch.zeros(24, 24, 24)
    for _ in range(6):
        cz, cy, cx = (torch.rand(3, generator=g) * 12 + 6).tolist()
        s = float(torch.rand(1, generator=g) * 2 + 1.5)
        vol += torch.exp(-((zz - cz) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
    R = uniform_rotations(6, 14)
    P = tfs.project_3d_to_2d(vol, R.clone())
    B = tfs.backproject_2d_to_3d(P, R.clone())
    save("project_backproject_24", volume=vol.numpy(), rotations=R.numpy(), projections=P.numpy(),
         backprojection=B.numpy())

    # G9: torch-autograd gradients at toy size (float64 leaf rotations).
    g = torch.Generator().manual_seed(15)
    v = torch.randn(8, 8, 8, generator=g)
    V = vol_rfft(v).detach().requires_grad_(True)
    R64 = uniform_rotations(3, 16).to(torch.float64).requires_grad_(True)
    T = torch.complex(torch.randn(3, 8, 5, generator=g), torch.randn(3, 8, 5, generator=g))
    S = ext(V, (8, 8, 8), R64)
    loss = torch.real(torch.sum(torch.conj(T) * S))
    loss.backward()
    ext_gV, ext_gR = V.grad.numpy().copy(), R64.grad.numpy().copy()
    D_ = torch.complex(torch.randn(3, 8, 5, generator=g), torch.randn(3, 8, 5, generator=g))
    D_ = D_.requires_grad_(True)
    R64b = uniform_rotations(3, 17).to(torch.float64).requires_grad_(True)
    TV = torch.complex(torch.randn(8, 8, 5, generator=g), torch.randn(8, 8, 5, generator=g))
    TW = torch.randn(8, 8, 5, generator=g)
    Vi, Wi = ins(D_, (8, 8, 8), R64b)
    loss = torch.real(torch.sum(torch.conj(TV) * Vi)) + torch.sum(TW * Wi)
    loss.backward()
    save("autograd_8", volume_rfft=V.detach().numpy(), ext_rotations=R64.detach().numpy(),
         ext_target=T.numpy(), ext_slices=S.detach().numpy(), ext_grad_volume=ext_gV,
         ext_grad_rotations=ext_gR, ins_slices=D_.detach().numpy(),
         ins_rotations=R64b.detach().numpy(), ins_target_volume=TV.numpy(),
         ins_target_weights=TW.numpy(), ins_grad_slices=D_.grad.numpy(),
         ins_grad_rotations=R64b.grad.numpy())

    # G10: in-plane 180-degree pairs with points on the x'=0 boundary, and the
    # rotations of the reference's insertion Friedel test.
    pairs = [
        ([0, 0, 0], [0, 0, 180]), ([0, 0, 0], [180, 0, 0]), ([0, 20, 0], [0, 20, 180]),
        ([0, 20, 90], [0, 20, 270]), ([0, 60, 0], [0, 60, 180]), ([90, 90, 0], [90, 90, 180]),
        ([45, 0, 0], [45, 0, 180]), ([45, 90, 0], [45, 90, 180]),
    ]
    g = torch.Generator().manual_seed(18)
    v = torch.randn(32, 32, 32, generator=g)
    V = vol_rfft(v)
    left = np.stack([zyz_matrix(a) for a, _ in pairs])
    right = np.stack([zyz_matrix(b) for _, b in pairs])
    SL = ext(V, (32, 32, 32), torch.from_numpy(left))
    SR = ext(V, (32, 32, 32), torch.from_numpy(right))
    ry90 = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], dtype=np.float64)
    rz90 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], dtype=np.float64)
    fried = np.concatenate([np.stack([zyz_matrix([0, 0, 0]), zyz_matrix([0, 0, 180])]),
                            ry90[None], rz90[None]])
    Sf = ext(V, (32, 32, 32), torch.from_numpy(fried))
    Vf, Wf = ins(Sf, (32, 32, 32), torch.from_numpy(fried))
    save("tiebreak_32", volume_rfft=V.numpy(), left=left, right=right, slices_left=SL.numpy(),
         slices_right=SR.numpy(), friedel_rotations=fried, friedel_slices=Sf.numpy(),
         friedel_volume=Vf.numpy(), friedel_weights=Wf.numpy())


def rotations_2d(n, seed):
    """(n, 2, 2) xy rotation matrices at uniform angles, plus the exact
    0/90/180/270-degree ones first (float32)."""
    import torch

    g = torch.Generator().manual_seed(seed)
    theta = torch.rand(n, generator=g, dtype=torch.float64) * 2 * np.pi
    c, s = torch.cos(theta), torch.sin(theta)
    R = torch.stack([c, -s, s, c], dim=-1).reshape(n, 2, 2)
    exact = torch.tensor([[[1.0, 0.0], [0.0, 1.0]], [[0.0, -1.0], [1.0, 0.0]],
                          [[-1.0, 0.0], [0.0, -1.0]], [[0.0, 1.0], [-1.0, 0.0]]], dtype=torch.float64)
    return torch.cat([exact, R]).to(torch.float32)


def image_rfft_2d(img):
    import torch

    f = torch.fft.fftshift(img, dim=(-2, -1))
    f = torch.fft.rfftn(f, dim=(-2, -1))
    return torch.fft.fftshift(f, dim=(-2,))


def main_2d(root):
    """The 2-D analogues: central lines of 2-D rffts, project_2d_to_1d, and
    the 2-D sampler / splat (TFS/slice_extraction/_extract_central_slices_rfft_2d.py,
    TFS/project.py:180-260, TII/image_interpolation_2d.py)."""
    import torch

    torch.set_num_threads(8)
    tfs, tii = _import_reference(root)

    def save(name, **arrays):
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
        print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")

    # G11: central lines of a 32x32 and an odd 33x33 N(0,1) image, and of a
    # 16x16 image with image_shape (32, 32).
    out = {}
    for key, n, shape, seed in (("even", 32, (32, 32), 21), ("odd", 33, (33, 33), 22),
                                ("shape2x", 16, (32, 32), 23)):
        g = torch.Generator().manual_seed(seed)
        F = image_rfft_2d(torch.randn(n, n, generator=g))
        R = rotations_2d(20, seed + 100)
        out[f"{key}_image_rfft"] = F.numpy()
        out[f"{key}_rotations"] = R.numpy()
        out[f"{key}_lines"] = tfs.extract_central_slices_rfft_2d(F, shape, R.clone()).numpy()
        out[f"{key}_lines_yx"] = tfs.extract_central_slices_rfft_2d(F, shape, R.clone(),
                                                                    yx_matrices=True).numpy()
    # G12: project_2d_to_1d of a 24x24 blob image, pad 2 and pad 1.
    g = torch.Generator().manual_seed(24)
    yy, xx = torch.meshgrid(torch.arange(24.0), torch.arange(24.0), indexing="ij")
    img = torch.zeros(24, 24)
    for _ in range(4):
        cy, cx = (torch.rand(2, generator=g) * 12 + 6).tolist()
        s = float(torch.rand(1, generator=g) * 2 + 1.5)
        img += torch.exp(-((yy - cy) ** 2 + (xx - cx) ** 2) / (2 * s * s))
    R = rotations_2d(8, 124).reshape(2, 6, 2, 2)
    out["proj_image"] = img.numpy()
    out["proj_rotations"] = R.numpy()
    out["proj_pad2"] = tfs.project_2d_to_1d(img, R.clone()).numpy()
    out["proj_pad1"] = tfs.project_2d_to_1d(img, R.clone(), pad_factor=1.0).numpy()
    save("extract_2d", **out)

    # G13: generic 2-D sampler, real / complex / multichannel, bilinear /
    # nearest / bicubic, coordinates partly out of bounds, .5 ties, integers.
    g = torch.Generator().manual_seed(25)
    img_r = torch.randn(12, 10, generator=g)
    img_c = torch.complex(torch.randn(12, 10, generator=g), torch.randn(12, 10, generator=g))
    img_m = torch.randn(3, 12, 10, generator=g)
    hi = torch.tensor([12.0, 10.0])
    crd = torch.rand(400, 2, generator=g) * (hi + 1.0) - 1.0
    ties = torch.randint(0, 9, (100, 2), generator=g).float() + 0.5
    ints = torch.randint(0, 10, (50, 2), generator=g).float()
    crd = torch.cat([crd, ties, ints]).reshape(11, 50, 2)
    out = {}
    for mode in ("bilinear", "nearest", "bicubic"):
        out[f"real_{mode}"] = tii.sample_image_2d(img_r, crd, interpolation=mode).numpy()
        out[f"complex_{mode}"] = tii.sample_image_2d(img_c, crd, interpolation=mode).numpy()
        out[f"multi_{mode}"] = tii.sample_image_2d(img_m, crd, interpolation=mode).numpy()
    save("sample_image_2d", image_real=img_r.numpy(), image_complex=img_c.numpy(),
         image_multi=img_m.numpy(), coordinates=crd.numpy(), **out)

    # G14: generic 2-D splat, bilinear / nearest, two accumulating calls.
    g = torch.Generator().manual_seed(26)
    crd1 = torch.cat([torch.rand(300, 2, generator=g) * (hi + 1.0) - 1.0, ties[:50]])
    crd2 = torch.rand(200, 2, generator=g) * (hi + 1.0) - 1.0
    val_r1, val_r2 = torch.randn(350, generator=g), torch.randn(200, generator=g)
    val_c1 = torch.complex(torch.randn(350, generator=g), torch.randn(350, generator=g))
    val_c2 = torch.complex(torch.randn(200, generator=g), torch.randn(200, generator=g))
    val_m1, val_m2 = torch.randn(350, 2, generator=g), torch.randn(200, 2, generator=g)
    out = {}
    for mode in ("bilinear", "nearest"):
        for key, v1, v2, img0 in (("real", val_r1, val_r2, torch.zeros(12, 10)),
                                  ("complex", val_c1, val_c2, torch.zeros(12, 10, dtype=torch.complex64)),
                                  ("multi", val_m1, val_m2, torch.zeros(2, 12, 10))):
            im, w = tii.insert_into_image_2d(v1, crd1, img0, interpolation=mode)
            im, w = tii.insert_into_image_2d(v2, crd2, im, w, interpolation=mode)
            out[f"{key}_{mode}_image"], out[f"{key}_{mode}_weights"] = im.numpy(), w.numpy()
    save("insert_into_image_2d", coords1=crd1.numpy(), coords2=crd2.numpy(),
         val_r1=val_r1.numpy(), val_r2=val_r2.numpy(), val_c1=val_c1.numpy(),
         val_c2=val_c2.numpy(), val_m1=val_m1.numpy(), val_m2=val_m2.numpy(), **out)


def affine_matrices_3d(n, seed, shape):
    """(n, 4, 4) float32 xyz homogeneous matrices: a rotation about the centre
    with a random scale and shift, plus a pure half-voxel translation first."""
    import torch
    from scipy.spatial.transform import Rotation

    rng = np.random.default_rng(seed)
    c = (np.array(shape[::-1], dtype=np.float64) - 1) / 2  # xyz centre
    out = [np.eye(4)]
    out[0][:3, 3] = [0.5, -0.5, 0.5]
    for k in range(n - 1):
        A = Rotation.random(random_state=seed + k).as_matrix() * rng.uniform(0.8, 1.2)
        t = c - A @ c + rng.uniform(-1.5, 1.5, 3)
        M = np.eye(4)
        M[:3, :3], M[:3, 3] = A, t
        out.append(M)
    return torch.from_numpy(np.stack(out)).to(torch.float32)


def affine_matrices_2d(n, seed, shape):
    import torch

    rng = np.random.default_rng(seed)
    c = (np.array(shape[::-1], dtype=np.float64) - 1) / 2
    out = [np.eye(3)]
    out[0][:2, 2] = [0.5, -0.5]
    for _ in range(n - 1):
        a = rng.uniform(0, 2 * np.pi)
        A = np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]]) * rng.uniform(0.8, 1.2)
        M = np.eye(3)
        M[:2, :2], M[:2, 2] = A, c - A @ c + rng.uniform(-1.5, 1.5, 2)
        out.append(M)
    return torch.from_numpy(np.stack(out)).to(torch.float32)


def main_transform(root):
    """Affine warps of torch-transform-image (transforms_3d.py, transforms_2d.py)."""
    import torch

    torch.set_num_threads(8)
    _import_reference(root)
    import torch_transform_image as tti  # noqa: E402

    def save(name, **arrays):
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
        print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")

    out = {}
    g = torch.Generator().manual_seed(31)
    imgs3 = {"real": torch.randn(12, 10, 9, generator=g),
             "complex": torch.complex(torch.randn(12, 10, 9, generator=g), torch.randn(12, 10, 9, generator=g)),
             "multi": torch.randn(2, 12, 10, 9, generator=g)}
    M3 = affine_matrices_3d(4, 32, (12, 10, 9)).reshape(4, 1, 1, 1, 4, 4)  # broadcast over (d, h, w)
    for key, img in imgs3.items():
        out[f"img3_{key}"] = img.numpy()
        for mode in ("trilinear", "nearest"):
            out[f"a3_{key}_{mode}"] = tti.affine_transform_image_3d(img, M3.clone(), mode).numpy()
            out[f"a3_{key}_{mode}_zyx_out"] = tti.affine_transform_image_3d(
                img, M3.clone(), mode, output_shape=(8, 7, 6), zyx_matrices=True).numpy()
    out["M3"] = M3.numpy()
    imgs2 = {"real": torch.randn(14, 11, generator=g),
             "complex": torch.complex(torch.randn(14, 11, generator=g), torch.randn(14, 11, generator=g)),
             "multi": torch.randn(3, 14, 11, generator=g)}
    M2 = affine_matrices_2d(5, 33, (14, 11)).reshape(5, 1, 1, 3, 3)
    for key, img in imgs2.items():
        out[f"img2_{key}"] = img.numpy()
        for mode in ("bilinear", "nearest", "bicubic"):
            out[f"a2_{key}_{mode}"] = tti.affine_transform_image_2d(img, M2.clone(), mode).numpy()
            out[f"a2_{key}_{mode}_yx_out"] = tti.affine_transform_image_2d(
                img, M2.clone(), mode, output_shape=(9, 8), yx_matrices=True).numpy()
    out["M2"] = M2.numpy()
    img = imgs2["real"]
    for mode in ("bilinear", "bicubic"):
        out[f"rts_{mode}"] = tti.rotate_then_shift_image_2d(img, rotate=30.0, shift=(2.5, -1.5),
                                                            interpolation=mode).numpy()
        out[f"str_{mode}"] = tti.shift_then_rotate_image_2d(img, rotate=30.0, shift=(2.5, -1.5),
                                                            interpolation=mode).numpy()
    out["shift_only"] = tti.rotate_then_shift_image_2d(img, shift=(1.0, 2.0), interpolation="bilinear").numpy()
    # one matrix per output column: the matrices broadcast against the grid
    Mcol = affine_matrices_2d(11, 34, (14, 11))
    out["Mcol"] = Mcol.numpy()
    out["a2_percol"] = tti.affine_transform_image_2d(img, Mcol.clone(), "bilinear").numpy()
    out["a3_single"] = tti.affine_transform_image_3d(imgs3["real"], M3[1, 0, 0, 0].clone(), "trilinear").numpy()
    save("transform_image", **out)

    # G16: subpixel crops (torch-subpixel-crop), real and rfft outputs
    import torch_subpixel_crop as tsc  # noqa: E402

    out = {}
    g = torch.Generator().manual_seed(35)
    vol = torch.randn(20, 18, 16, generator=g)
    pos3 = torch.rand(2, 3, 3, generator=g) * torch.tensor([20.0, 18.0, 16.0]) - 1.0  # some patches cross the edge
    pos3[0, 0] = torch.tensor([9.5, 8.5, 7.25])  # round-half-to-even ties
    mask3 = torch.rand(8, 8, 8, generator=g)
    out["vol"], out["pos3"], out["mask3"] = vol.numpy(), pos3.numpy(), mask3.numpy()
    out["c3"] = tsc.subpixel_crop_3d(vol, pos3, 8).numpy()
    out["c3_mask"] = tsc.subpixel_crop_3d(vol, pos3, 8, mask=mask3).numpy()
    out["c3_rfft"] = tsc.subpixel_crop_3d(vol, pos3, 8, return_rfft=True).numpy()
    out["c3_rfft_dec"] = tsc.subpixel_crop_3d(vol, pos3, 8, return_rfft=True, decenter=True).numpy()
    imgs = torch.randn(3, 30, 26, generator=g)
    pos2 = torch.rand(4, 3, 2, generator=g) * torch.tensor([30.0, 26.0]) - 1.0
    pos2[0, 0] = torch.tensor([10.5, 11.5])
    mask2 = torch.rand(10, 10, generator=g)
    out["imgs"], out["pos2"], out["mask2"] = imgs.numpy(), pos2.numpy(), mask2.numpy()
    out["c2"] = tsc.subpixel_crop_2d(imgs, pos2, 10).numpy()
    out["c2_mask"] = tsc.subpixel_crop_2d(imgs, pos2, 10, mask=mask2).numpy()
    out["c2_rfft_dec"] = tsc.subpixel_crop_2d(imgs, pos2, 10, return_rfft=True, decenter=True).numpy()
    out["c2_single"] = tsc.subpixel_crop_2d(imgs[1], pos2[:, 1], 9).numpy()  # unbatched, odd sidelength
    save("subpixel_crop", **out)


def main_mc_grad(root):
    """Multichannel extraction, insertion and projection
    (slice_extraction/_extract_central_slices_rfft_3d.py:225-367,
    slice_insertion/_insert_central_slices_rfft_3d.py:109-210,
    project.py:96-177), and the torch-autograd gradients of sample_image_3d
    (image_interpolation_3d.py:11-101) and affine_transform_image_3d
    (transforms_3d.py:10-40) with respect to the image, the coordinates and
    the matrices."""
    import torch

    torch.set_num_threads(8)
    tfs, tii = _import_reference(root)
    import torch_transform_image as tti  # noqa: E402

    def save(name, **arrays):
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
        print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB)")

    # G17: multichannel operators, three channels (an odd count)
    g = torch.Generator().manual_seed(41)
    vol = torch.randn(3, 16, 16, 16, generator=g)
    V = vol_rfft(vol)
    R = uniform_rotations(5, 42)
    out = {"volume_rfft": V.numpy(), "rotations": R.numpy()}
    out["slices"] = tfs.extract_central_slices_rfft_3d_multichannel(V, (16, 16, 16), R.clone()).numpy()
    out["slices_fmax"] = tfs.extract_central_slices_rfft_3d_multichannel(
        V, (16, 16, 16), R.clone(), fftfreq_max=0.3).numpy()
    S = torch.complex(torch.randn(5, 3, 16, 9, generator=g), torch.randn(5, 3, 16, 9, generator=g))
    Vi, Wi = tfs.insert_central_slices_rfft_3d_multichannel(S, (16, 16, 16), R.clone())
    out["ins_slices"], out["ins_volume"], out["ins_weights"] = S.numpy(), Vi.numpy(), Wi.numpy()
    Vf, Wf = tfs.insert_central_slices_rfft_3d_multichannel(S, (16, 16, 16), R.clone(), fftfreq_max=0.3)
    out["ins_volume_fmax"], out["ins_weights_fmax"] = Vf.numpy(), Wf.numpy()
    real = torch.randn(2, 12, 12, 12, generator=g)
    Rp = uniform_rotations(4, 43)
    out["proj_volume"], out["proj_rotations"] = real.numpy(), Rp.numpy()
    out["projections"] = tfs.project_3d_to_2d_multichannel(real, Rp.clone()).numpy()
    save("multichannel_16", **out)

    # G18: autograd of sample_image_3d and affine_transform_image_3d
    out = {}
    g = torch.Generator().manual_seed(44)
    shape = (7, 6, 8)
    crd = torch.rand(60, 3, generator=g) * (torch.tensor(shape, dtype=torch.float32) + 1.0) - 1.0
    crd[:6] = torch.tensor([[0.0, 2.5, 3.25], [6.0, 1.5, 2.0], [3.5, 0.0, 7.0], [2.25, 5.0, 4.5],
                            [1.0, 2.0, 3.0], [6.0, 5.0, 7.0]])  # on faces, edges and voxels
    out["coords"] = crd.numpy()
    imgs = {"real": torch.randn(*shape, generator=g),
            "complex": torch.complex(torch.randn(*shape, generator=g), torch.randn(*shape, generator=g)),
            "multi": torch.randn(2, *shape, generator=g)}
    for key, img in imgs.items():
        out[f"img_{key}"] = img.numpy()
        for mode in ("trilinear", "nearest"):
            im = img.clone().requires_grad_(True)
            c = crd.clone().requires_grad_(True)
            smp = tii.sample_image_3d(im, c, interpolation=mode)
            tgt = torch.randn(smp.shape, generator=g, dtype=torch.float32)
            if smp.is_complex():
                tgt = torch.complex(tgt, torch.randn(smp.shape, generator=g))
                loss = torch.real(torch.sum(torch.conj(tgt) * smp))
            else:
                loss = torch.sum(tgt * smp)
            loss.backward()
            out[f"{key}_{mode}_target"] = tgt.numpy()
            out[f"{key}_{mode}_samples"] = smp.detach().numpy()
            out[f"{key}_{mode}_grad_image"] = im.grad.numpy()
            out[f"{key}_{mode}_grad_coords"] = c.grad.numpy()
    img = torch.randn(2, 9, 8, 7, generator=g)
    M = affine_matrices_3d(3, 45, (9, 8, 7)).reshape(3, 1, 1, 1, 4, 4)
    im = img.clone().requires_grad_(True)
    Mg = M.clone().requires_grad_(True)
    warped = tti.affine_transform_image_3d(im, Mg, "trilinear")
    tgt = torch.randn(warped.shape, generator=g)
    torch.sum(tgt * warped).backward()
    out.update(affine_image=img.numpy(), affine_matrices=M.numpy(), affine_target=tgt.numpy(),
               affine_warped=warped.detach().numpy(), affine_grad_image=im.grad.numpy(),
               affine_grad_matrices=Mg.grad.numpy())
    save("sample_autograd", **out)


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    root = args[0] if args else "/root/reference"
    if "--2d" in sys.argv:  # only the fixtures of the 2-D analogues
        main_2d(root)
    elif "--transform" in sys.argv:  # only the affine-warp fixtures
        main_transform(root)
    elif "--mc-grad" in sys.argv:  # only the multichannel and sampler-autograd fixtures
        main_mc_grad(root)
    else:
        main(root)
        main_2d(root)
        main_transform(root)
        main_mc_grad(root)
