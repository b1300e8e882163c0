"""Workload-shaped GPU checks beyond the unit parity tests.

* Differentiation: a CTF-weighted squared-error loss on projections of a
  128^3 blob volume (pad 2), gradients w.r.t. the volume and the rotation
  matrices against FP64 central finite differences of the same loss built on
  the FP64 oracle.
* Tomogram backprojection: a 61-tilt single-axis series of 512^2 images into a
  (512, 512, 257) rfft volume; the insertion against the FP64 C oracle, and a
  few steps of gradient descent on per-tilt shifts through the backward.
"""

import math

import numpy as np
import pytest
import torch

import teamtomo_amd as tt
from conftest import rel_l2
from oracle import c_oracle as C
from oracle import fourier_slice as O
from teamtomo_amd.synthetic import blob_volume, uniform_rotations

pytestmark = pytest.mark.gpu


# ------------------------------------------------------------------ the CTF
def ctf_rfft(n, pixel_size, defocus_u, defocus_v, astig_angle, kv=300.0, cs_mm=2.7, amp=0.07,
             xp=np):
    """Phase-contrast transfer function on the rfft grid (n, n//2+1) of an n^2
    image, one per row of the defocus arrays (Angstrom, degrees)."""
    volts = kv * 1e3
    lam = 12.2643 / math.sqrt(volts * (1.0 + 0.978466e-6 * volts))  # Angstrom
    cs = cs_mm * 1e7
    ky = np.fft.fftfreq(n) / pixel_size
    kx = np.fft.rfftfreq(n) / pixel_size
    k2 = (ky[:, None] ** 2 + kx[None, :] ** 2)
    ang = np.arctan2(ky[:, None] + 0.0 * kx[None, :], kx[None, :] + 0.0 * ky[:, None])
    du = np.asarray(defocus_u, dtype=np.float64)[:, None, None]
    dv = np.asarray(defocus_v, dtype=np.float64)[:, None, None]
    aa = np.deg2rad(np.asarray(astig_angle, dtype=np.float64))[:, None, None]
    df = 0.5 * (du + dv) + 0.5 * (du - dv) * np.cos(2.0 * (ang[None] - aa))
    chi = math.pi * lam * df * k2[None] - 0.5 * math.pi * cs * lam ** 3 * (k2[None] ** 2)
    c = -(math.sqrt(1.0 - amp * amp) * np.sin(chi) + amp * np.cos(chi))
    return c


def test_ctf_loss_gradients_128cubed_512poses_vs_finite_differences(device):
    """128^3 blob volume, pad 2 (256^3 inside), 512 poses, CTF at
    300 kV / Cs 2.7 mm / 0.07 / 1.5 A/px, defocus U(1,3) um, astigmatism
    U(0,0.1) um; loss = sum (ctf * proj - target)^2 with target = true
    projections + noise at SNR 0.1.  Gradients w.r.t. 8 poses' matrices (all
    9 entries) and w.r.t. the volume along 2 random directions against FP64
    central differences."""
    n, n_pose = 128, 512
    rng = np.random.default_rng(30)
    vol_true = blob_volume(n, 60, seed=31, device=device)
    vol = vol_true + 0.05 * torch.from_numpy(rng.standard_normal((n, n, n)).astype(np.float32)).to(device)
    R = uniform_rotations(n_pose, 32, device=device)
    dU = rng.uniform(1e4, 3e4, n_pose)
    dV = dU + rng.uniform(0.0, 1e3, n_pose)
    ang = rng.uniform(0.0, 180.0, n_pose)
    ctf = ctf_rfft(n, 1.5, dU, dV, ang)
    ctf_t = torch.from_numpy(ctf.astype(np.float32)).to(device)

    def model(v, Rm):
        p = tt.project_3d_to_2d(v, Rm, pad_factor=2.0)
        return torch.fft.irfftn(torch.fft.rfftn(p, dim=(-2, -1)) * ctf_t, s=(n, n), dim=(-2, -1))

    with torch.no_grad():
        clean = model(vol_true, R)
        sigma = math.sqrt(float(clean.var()) / 0.1)
        target = clean + sigma * torch.from_numpy(rng.standard_normal(clean.shape).astype(np.float32)).to(device)
    v = vol.clone().requires_grad_(True)
    Rg = R.clone().to(torch.float64).requires_grad_(True)
    loss = torch.sum((model(v, Rg) - target) ** 2)
    loss.backward()

    # FP64 restatement: the oracle's projection pipeline with its DFT cached
    vol64 = vol.double().cpu().numpy()
    tgt64 = target.double().cpu().numpy()
    Rn = R.double().cpu().numpy()
    p = n // 2

    def dft_of(vreal):
        vp = np.pad(vreal, p)
        return O.volume_rfft(vp * O._sinc2_grid(vp.shape))

    def loss64(dft, Rz, idx, exact_rot=False):
        # Rz: zyx matrices; the FP64 NumPy oracle keeps FP64 rotations (for
        # differences in R), the C oracle takes them as float32
        s = O.extract_central_slices(dft, (2 * n,) * 3, Rz) if exact_rot else C.extract(dft, (2 * n,) * 3, Rz)
        s = np.fft.ifftshift(s, axes=(-2,))
        proj = np.fft.ifftshift(np.fft.irfftn(s, axes=(-2, -1)), axes=(-2, -1))[..., p:-p, p:-p]
        out = np.fft.irfftn(np.fft.rfftn(proj, axes=(-2, -1)) * ctf[idx], s=(n, n), axes=(-2, -1))
        return float(np.sum((out - tgt64[idx]) ** 2))

    dft0 = dft_of(vol64)
    Rz = O.prepare_rotations(Rn, False, 1e-8).astype(np.float64)
    # rotation gradients: each pose's loss term depends on its own matrix only
    poses = [0, 57, 130, 222, 301, 399, 451, 511]
    # Trilinear interpolation has kinks where a sample crosses a cell face (or
    # the volume's bounds, or the Friedel half changes): a central difference
    # across one is wrong.  The step starts small (coordinates up to 128
    # voxels move by 128 h) and is cut until no pixel of the pose changes its
    # cell, half or bounds between R - h E and R + h E, so the loss is a
    # polynomial of the entry over the whole stencil.
    def cells(Rz1):
        fr = O._rotated_frequencies(Rz1, 2 * n, 2 * n)
        conj = O._extract_conjugate_mask(fr, Rz1, 1e-8)
        c = O._to_array_coords(np.where(conj[..., None], -fr, fr), (2 * n,) * 3)
        ok = O._inside(c, (2 * n, 2 * n, n + 1))
        return np.floor(c), conj, ok

    fd_R = np.zeros((len(poses), 3, 3))
    steps = []
    for a, k in enumerate(poses):
        for i in range(3):
            for j in range(3):
                h = 1e-8
                while True:
                    Rp, Rm = Rz[k:k + 1].copy(), Rz[k:k + 1].copy()
                    Rp[0, i, j] += h
                    Rm[0, i, j] -= h
                    cp, cm = cells(Rp), cells(Rm)
                    if all(np.array_equal(u, v) for u, v in zip(cp, cm)) or h < 2e-10:
                        break
                    h /= 4
                steps.append(h)
                fd_R[a, i, j] = (loss64(dft0, Rp, [k], True) - loss64(dft0, Rm, [k], True)) / (2 * h)
    print(f"  finite-difference steps: {sorted(set(steps))}")
    fd_R = O.rotation_grad_to_xyz(fd_R, Rn[poses], False, 1e-8)
    g_R = Rg.grad.detach().cpu().numpy()[poses]
    err_R = rel_l2(g_R, fd_R)
    for a in range(len(poses)):
        print(f"  pose {poses[a]}: rel_l2 {rel_l2(g_R[a], fd_R[a]):.2e}  |g| {np.linalg.norm(g_R[a]):.3e}")
    # volume gradient along random directions: the loss is quadratic in the
    # volume, so the central difference is exact up to rounding for any step
    all_idx = list(range(n_pose))
    err_V = []
    for seed in (33, 34):
        d = np.random.default_rng(seed).standard_normal((n, n, n))
        hv = 1e-2
        fd = (loss64(dft_of(vol64 + hv * d), Rz, all_idx) - loss64(dft_of(vol64 - hv * d), Rz, all_idx)) / (2 * hv)
        g = float(np.sum(v.grad.double().cpu().numpy() * d))
        err_V.append(abs(g - fd) / abs(fd))
    print(f"128^3 pad 2 x 512 loss {float(loss.detach()):.6e}: grad R rel_l2 vs FP64 FD {err_R:.2e}; "
          f"grad V directional rel err {err_V[0]:.2e}, {err_V[1]:.2e}")
    assert err_R < 1e-4
    assert max(err_V) < 1e-4


# ----------------------------------------------------- tomogram backprojection
def _tilt_rotations(tilts_deg):
    t = np.deg2rad(np.asarray(tilts_deg, dtype=np.float64))
    R = np.zeros((len(t), 3, 3))
    R[:, 0, 0] = np.cos(t)
    R[:, 0, 2] = np.sin(t)
    R[:, 1, 1] = 1.0
    R[:, 2, 0] = -np.sin(t)
    R[:, 2, 2] = np.cos(t)
    return R.astype(np.float32)


def _phase_shift(images, shifts):
    """Shift 2-D images by (dy, dx) pixels with a Fourier phase ramp."""
    n = images.shape[-1]
    ky = torch.fft.fftfreq(n, device=images.device, dtype=torch.float64)[:, None]
    kx = torch.fft.rfftfreq(n, device=images.device, dtype=torch.float64)[None, :]
    s = shifts.to(torch.float64)
    ph = -2.0 * math.pi * (ky[None] * s[:, 0, None, None] + kx[None] * s[:, 1, None, None])
    ramp = torch.polar(torch.ones_like(ph), ph).to(torch.complex64)
    return torch.fft.irfftn(torch.fft.rfftn(images, dim=(-2, -1)) * ramp, s=(n, n), dim=(-2, -1))


def test_61tilt_512cubed_backprojection_and_shift_fit(device):
    """61 tilts -60..60 step 2, 512^2 float32 images = N(0,1) + 500 Gaussian
    features (projections of 500 seeded 3-D blobs), per-tilt shifts
    U(-5,5) px applied as phase ramps (torch, FP64).  The insertion into
    (512,512,257) is checked against the FP64 C oracle; backprojection runs
    with pad_factor=1; 40 Adam steps on the per-tilt shifts through the
    backward (fourier_shift_image_2d, then backproject_2d_to_3d),
    fitting the low-passed backprojection of the unshifted series, must cut
    the loss tenfold and the shift error fourfold."""
    n = 512
    tilts = np.arange(-60, 61, 2)
    R = torch.from_numpy(_tilt_rotations(tilts)).to(device)
    gen = torch.Generator().manual_seed(50)
    feat = blob_volume(n, 500, seed=51, sigma=(1.5, 3.0), noise=0.0, device=device)
    with torch.no_grad():
        clean = tt.project_3d_to_2d(feat, R, pad_factor=1.0)
        noise = torch.randn(clean.shape, generator=gen).to(device)
        shifts_true = (torch.rand(len(tilts), 2, generator=gen, dtype=torch.float64) * 10.0 - 5.0)
        images = _phase_shift(clean + noise, shifts_true.to(device))
    del feat

    # insertion parity at full size: the backprojection's own input
    # (the default path for this call is the binned scatter; the brick engine
    # is checked too, and the binned result must repeat bit for bit)
    with torch.no_grad():
        x = torch.fft.fftshift(images, dim=(-2, -1))
        x = torch.fft.rfftn(x, dim=(-2, -1))
        x = torch.fft.fftshift(x, dim=(-2,))
        dft, w = tt.insert_central_slices_rfft_3d(x, (n, n, n), R)
        dft2, w2 = tt.insert_central_slices_rfft_3d(x, (n, n, n), R)
        assert torch.equal(dft, dft2) and torch.equal(w, w2)
        del dft2, w2
        tt.set_scatter_path("brick")
        try:
            dft_b, w_b = tt.insert_central_slices_rfft_3d(x, (n, n, n), R)
        finally:
            tt.set_scatter_path("auto")
    Rz = O.prepare_rotations(R.cpu().numpy(), False, None)
    ref_v, ref_w = C.insert(x.cpu().numpy(), (n, n, n), Rz)
    ev, ew = rel_l2(dft.cpu().numpy(), ref_v), rel_l2(w.cpu().numpy(), ref_w)
    evb, ewb = rel_l2(dft_b.cpu().numpy(), ref_v), rel_l2(w_b.cpu().numpy(), ref_w)
    print(f"61-tilt insert (512,512,257): rel_l2 volume {ev:.2e}, weights {ew:.2e} "
          f"(brick engine: {evb:.2e}, {ewb:.2e})")
    assert ev < 1e-5 and ew < 1e-5 and evb < 1e-5 and ewb < 1e-5
    del dft, w, dft_b, w_b, ref_v, ref_w, x

    rec = tt.backproject_2d_to_3d(images, R, pad_factor=1.0)
    assert rec.shape == (n, n, n) and bool(torch.isfinite(rec).all())
    del rec

    # shift fit: the per-tilt shifts that make the low-passed backprojection
    # match the one of the unshifted series (least squares, through the backward)
    fm = 0.05
    with torch.no_grad():
        target = tt.backproject_2d_to_3d(clean + noise, R, pad_factor=1.0, fftfreq_max=fm)
    s = torch.zeros(len(tilts), 2, dtype=torch.float64, device=device, requires_grad=True)
    opt = torch.optim.Adam([s], lr=0.3)
    losses = []
    err0 = float(torch.sqrt(torch.mean((s.detach().cpu() - shifts_true) ** 2)))
    for _ in range(40):
        opt.zero_grad()
        # the shifts' gradient comes from the Fourier-shift kernel's backward
        moved = tt.fourier_shift_image_2d(images, -s.float())
        rec = tt.backproject_2d_to_3d(moved, R, pad_factor=1.0, fftfreq_max=fm)
        loss = torch.sum((rec - target).double() ** 2)
        loss.backward()
        opt.step()
        losses.append(float(loss.detach()))
        del rec, loss, moved
    err1 = float(torch.sqrt(torch.mean((s.detach().cpu() - shifts_true) ** 2)))
    print(f"61-tilt shift fit: loss {losses[0]:.4e} -> {losses[-1]:.4e}, shift error rms {err0:.3f} -> {err1:.3f} px")
    assert losses[-1] < 0.1 * losses[0]
    assert err1 < 0.25 * err0
