"""FP64 NumPy restatement of the reference Fourier-slice gather/scatter path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg, never by the ``teamtomo_amd`` package.

Conventions (all arrays zyx, C-contiguous):

* A 3-D rfft volume is ``fftshift(rfftn(fftshift(v)), axes=(-3, -2))``: DC at
  ``(D//2, H//2, 0)``.  A 2-D rfft slice has DC at ``(H//2, 0)``.
* ``R_zyx`` are the rotation matrices after the reference's float32 cast,
  optional near-zero clamp and xyz->zyx flip (``prepare_rotations``).  Every
  later step is carried out in float64 from those float32 values.
* Discrete decisions the reference makes in float32 on exactly representable
  inputs (the ``fftfreq_max`` disc, tolerance thresholds) are made in float32
  here too, so that they agree with the reference bit for bit.

Reference roots used in the citations below:
``TFS`` = packages/primitives/torch-fourier-slice/src/torch_fourier_slice,
``TII`` = packages/primitives/torch-image-interpolation/src/torch_image_interpolation,
``TGU`` = packages/primitives/torch-grid-utils/src/torch_grid_utils.

Pinning (what checks each function; see tests/test_oracle.py):
prepare_rotations, central_slice_frequencies, fftfreq_mask -- closed forms
and the golden fixtures; extract_central_slices -- R=I known answer, in-plane
180-degree conjugate pairs, x=0 Friedel line, golden fixtures; extract_adjoint
-- FP64 dot-product test against extract; extract_rotation_grad -- central
finite differences and golden torch-autograd gradients; insert_central_slices
-- known-answer splats, x=0 Friedel invariant, golden fixtures;
insert_adjoint / insert_rotation_grad -- dot-product test and finite
differences; sample_image_3d / insert_into_image_3d -- textbook trilinear
(affine reproduction), known answers from the reference tests, golden
fixtures; project_3d_to_2d / backproject_2d_to_3d -- rotation-centre and
central-slice-theorem tests, golden fixtures; the multichannel operators --
the golden fixtures ``multichannel_16`` and equality with the single-channel
functions applied per channel; sample_image_3d_grad -- golden torch-autograd
gradients (``sample_autograd``), the adjoint identity of the image gradient
and FP64 central differences of the coordinate gradient.
"""

from __future__ import annotations

import numpy as np

__all__ = [
    "rfft_width",
    "prepare_rotations",
    "central_slice_frequencies",
    "fftfreq_mask",
    "extract_central_slices",
    "extract_adjoint",
    "extract_rotation_grad",
    "insert_central_slices",
    "insert_adjoint",
    "insert_rotation_grad",
    "sample_image_3d",
    "sample_image_3d_grad",
    "insert_into_image_3d",
    "extract_central_slices_multichannel",
    "insert_central_slices_multichannel",
    "project_3d_to_2d_multichannel",
    "project_3d_to_2d",
    "backproject_2d_to_3d",
    "volume_rfft",
    "rotation_grad_to_xyz",
]


def rfft_width(w: int) -> int:
    """Last-dim length of an rfft: ``int(w/2 + 1)`` (TGU/fftfreq_grid.py:220-224)."""
    return int(w / 2 + 1)


def prepare_rotations(R, zyx_matrices: bool = False, rot_tolerance=None) -> np.ndarray:
    """Cast to float32, zero ``|R| < rot_tolerance`` and flip xyz->zyx.

    Extraction: TFS/slice_extraction/_extract_central_slices_rfft_3d.py:147-154.
    Insertion (no clamp): TFS/slice_insertion/_insert_central_slices_rfft_3d.py:16,66-67.
    The comparison is made in float32, as torch compares a float32 tensor
    with a Python float.  The caller's array is not modified.
    """
    R32 = np.array(R, dtype=np.float32, copy=True)
    if rot_tolerance is not None:
        R32[np.abs(R32) < np.float32(rot_tolerance)] = 0.0
    if not zyx_matrices:
        R32 = R32[..., ::-1, ::-1]
    return np.ascontiguousarray(R32)


def rotation_grad_to_xyz(g_zyx, R_input, zyx_matrices=False, rot_tolerance=None):
    """Chain a gradient w.r.t. ``R_zyx`` back to the caller's matrices.

    Inverse of ``prepare_rotations``: un-flip, and zero the entries that the
    near-zero clamp replaced by a constant (their gradient is 0 in torch's
    ``index_put_`` backward, TFS/slice_extraction/_extract_central_slices_rfft_3d.py:149-151).
    """
    g = np.array(g_zyx, dtype=np.float64, copy=True)
    if not zyx_matrices:
        g = g[..., ::-1, ::-1]
    if rot_tolerance is not None:
        R32 = np.asarray(R_input, dtype=np.float32)
        g = np.where(np.abs(R32) < np.float32(rot_tolerance), 0.0, g)
    return np.ascontiguousarray(g)


def central_slice_frequencies(H: int, W: int):
    """Frequencies of the central (z=0) rfft slice, float64.

    ``fy[i] = (i - H//2)/H`` (fftshifted fftfreq along rows) and
    ``fx[j] = j/W`` (rfftfreq), i in [0,H), j in [0, W//2+1).
    TFS/_grids/_central_slice_fftfreq_grid.py:6-32, TGU/fftfreq_grid.py:142-177,291-311.
    Returned with shapes (H, 1) and (1, Wr) for broadcasting.
    """
    Wr = rfft_width(W)
    fy = (np.arange(H, dtype=np.float64) - H // 2) / H
    fx = np.arange(Wr, dtype=np.float64) / W
    return fy[:, None], fx[None, :]


def _fftfreq_f32(n: int) -> np.ndarray:
    """torch.fft.fftfreq(n) in float32: arange values times float32(1/n)."""
    k = np.concatenate([np.arange(0, (n + 1) // 2), np.arange(-(n // 2), 0)])
    return (k.astype(np.float32) * np.float32(1.0 / n)).astype(np.float32)


def fftfreq_mask(H: int, W: int, fftfreq_max) -> np.ndarray:
    """Boolean (H, Wr) disc ``sqrt(fy^2 + fx^2) <= fftfreq_max`` evaluated in float32.

    TFS/slice_extraction/_extract_central_slices_rfft_3d.py:170-180 (and the
    same block in insertion :30-38).  The float32 grid values come from
    torch.fft.fftfreq / rfftfreq (TGU/fftfreq_grid.py:172-173).
    """
    Wr = rfft_width(W)
    if fftfreq_max is None:
        return np.ones((H, Wr), dtype=bool)
    fy = np.fft.fftshift(_fftfreq_f32(H))
    fx = (np.arange(Wr).astype(np.float32) * np.float32(1.0 / W)).astype(np.float32)
    r2 = (fy[:, None] * fy[:, None]).astype(np.float32) + (fx[None, :] * fx[None, :]).astype(np.float32)
    r = np.sqrt(r2.astype(np.float32)).astype(np.float32)
    return r <= np.float32(fftfreq_max)


def _full_scale(n: int) -> float:
    """Array coordinate per cycle/pixel on a full fftshifted axis: (n-1)/(1-1/n).

    TFS/_dft_utils.py:46-64 (equals n in exact arithmetic).
    """
    return (n - 1) / (1.0 - 1.0 / n)


def _rotated_frequencies(R_zyx, H, W):
    """zyx rotated frequencies of every slice pixel, float64, shape (n, H, Wr, 3)."""
    R = np.asarray(R_zyx, dtype=np.float64).reshape(-1, 3, 3)
    fy, fx = central_slice_frequencies(H, W)
    out = np.empty((R.shape[0], H, fx.shape[1], 3), dtype=np.float64)
    for d in range(3):
        out[..., d] = R[:, d, 1, None, None] * fy[None] + R[:, d, 2, None, None] * fx[None]
    return out


#: FP64 coordinates within this distance of an integer are snapped to it.
#: The reference computes coordinates in float32, which absorbs the 1e-16
#: noise of exact-geometry rotations (e.g. scipy's 90/180-degree matrices) and
#: lands exactly on integers / the volume faces; FP64 would land 1e-15 away
#: and flip the in-bounds test and the floor/ceil corner choice.  1e-9 voxel
#: is far below float32 resolution (>= 6e-8 relative), so random geometry is
#: unaffected.
SNAP_EPS = 1e-9


def _snap(c):
    r = np.rint(c)
    return np.where(np.abs(c - r) < SNAP_EPS, r, c)


def _to_array_coords(fr, shape_zyx):
    """cycles/pixel -> fftshifted rfft array coordinates (TFS/_dft_utils.py:5-72)."""
    D, H, W = shape_zyx
    Wr = rfft_width(W)
    c = np.empty_like(fr)
    c[..., 0] = (fr[..., 0] + 0.5) / (1.0 - 1.0 / D) * (D - 1)
    c[..., 1] = (fr[..., 1] + 0.5) / (1.0 - 1.0 / H) * (H - 1)
    c[..., 2] = (fr[..., 2] * 2) * (Wr - 1)
    return _snap(c)


def _coord_scales(shape_zyx):
    D, H, W = shape_zyx
    return np.array([_full_scale(D), _full_scale(H), 2.0 * (rfft_width(W) - 1)])


def _extract_conjugate_mask(fr, R_zyx, x_tolerance):
    """Which rotated points are read from the Friedel mate.

    TFS/slice_extraction/_extract_central_slices_rfft_3d.py:9-81,193-201:
    ``x' < -tol``; for ``|x'| < tol`` decide by the sign of R_zyx[2,2], or of
    R_zyx[2,1] when ``|R_zyx[2,2]| <= 10 tol``.  Thresholds are float32.
    """
    R = np.asarray(R_zyx, dtype=np.float64).reshape(-1, 3, 3)
    tol = float(np.float32(x_tolerance))
    tol10 = float(np.float32(x_tolerance * 10))
    x = fr[..., 2]
    conj = x < -tol
    near = np.abs(x) < tol
    r11 = R[:, 2, 2][:, None, None]
    r12 = R[:, 2, 1][:, None, None]
    tie = np.where(np.abs(r11) <= tol10, r12 < 0, r11 < 0)
    return conj | (near & tie)


def _trilinear_corners(c, shape):
    """floor / floor+1 corners and weights; corners past N-1 get weight 0.

    Returns list of 8 (index tuple, weight, valid) and the per-axis
    fractional parts.
    """
    i0 = np.floor(c).astype(np.int64)
    t = c - i0
    out = []
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                iz, iy, ix = i0[..., 0] + dz, i0[..., 1] + dy, i0[..., 2] + dx
                wz = t[..., 0] if dz else 1 - t[..., 0]
                wy = t[..., 1] if dy else 1 - t[..., 1]
                wx = t[..., 2] if dx else 1 - t[..., 2]
                valid = (iz < shape[0]) & (iy < shape[1]) & (ix < shape[2])
                out.append(((iz, iy, ix), wz * wy * wx, valid, (dz, dy, dx)))
    return out, i0, t


def _inside(c, shape):
    """All coordinates in [0, N-1] (TII/image_interpolation_3d.py:85-89)."""
    hi = np.array(shape, dtype=np.float64) - 1
    return np.all((c >= 0) & (c <= hi), axis=-1)


def _gather_trilinear(vol, c, ok):
    """Sample ``vol`` (D,H,Wr) at coordinates c (..., 3) where ``ok``; else 0."""
    shape = vol.shape
    out = np.zeros(c.shape[:-1], dtype=np.result_type(vol.dtype, np.float64))
    corners, _, _ = _trilinear_corners(np.where(ok[..., None], c, 0.0), shape)
    for (iz, iy, ix), w, valid, _ in corners:
        v = valid & ok
        out += np.where(v, w, 0.0) * vol[np.where(v, iz, 0), np.where(v, iy, 0), np.where(v, ix, 0)]
    return out


def extract_central_slices(volume, image_shape, R_zyx, fftfreq_max=None, x_tolerance=1e-8):
    """Central slices of an fftshifted rfft volume, one per rotation (float64).

    TFS/slice_extraction/_extract_central_slices_rfft_3d.py:84-222 with
    TFS/_dft_utils.py:5-72 and the trilinear sampler of
    TII/image_interpolation_3d.py:11-101 (torch grid_sample, border padding,
    align_corners=True, zero outside [0, N-1]).  ``image_shape`` drives the
    frequency grid and coordinate scale, the volume's own shape the bounds.
    Returns (n, H_i, Wr_i) complex128 (leading dims of R_zyx flattened).
    """
    vol = np.asarray(volume, dtype=np.complex128)
    Di, Hi, Wi = image_shape
    R = np.asarray(R_zyx, dtype=np.float64).reshape(-1, 3, 3)
    fr = _rotated_frequencies(R, Hi, Wi)
    conj = _extract_conjugate_mask(fr, R, x_tolerance)
    fr = np.where(conj[..., None], -fr, fr)
    c = _to_array_coords(fr, (Di, Hi, Wi))
    mask = fftfreq_mask(Hi, Wi, fftfreq_max)[None]
    ok = _inside(c, vol.shape) & mask
    s = _gather_trilinear(vol, c, ok)
    s = np.where(conj, np.conj(s), s)
    return np.where(ok, s, 0.0).astype(np.complex128)


def extract_adjoint(grad_slices, volume_shape_rfft, image_shape, R_zyx, fftfreq_max=None,
                    x_tolerance=1e-8):
    """Adjoint of ``extract_central_slices`` w.r.t. the volume (grad of volume).

    Same coordinates, conjugation and corner rule as the forward gather; no
    Friedel mirror and no weights (the linear map whose transpose torch's
    grid_sample backward applies, TII/image_interpolation_3d.py:69-75).
    ``volume_shape_rfft`` is the (D, H, Wr) shape of the volume.
    """
    g = np.asarray(grad_slices, dtype=np.complex128)
    Di, Hi, Wi = image_shape
    R = np.asarray(R_zyx, dtype=np.float64).reshape(-1, 3, 3)
    g = g.reshape(R.shape[0], Hi, rfft_width(Wi))
    fr = _rotated_frequencies(R, Hi, Wi)
    conj = _extract_conjugate_mask(fr, R, x_tolerance)
    fr = np.where(conj[..., None], -fr, fr)
    c = _to_array_coords(fr, (Di, Hi, Wi))
    mask = fftfreq_mask(Hi, Wi, fftfreq_max)[None]
    ok = _inside(c, volume_shape_rfft) & mask
    gk = np.where(conj, np.conj(g), g)
    gV = np.zeros(volume_shape_rfft, dtype=np.complex128)
    corners, _, _ = _trilinear_corners(np.where(ok[..., None], c, 0.0), volume_shape_rfft)
    for (iz, iy, ix), w, valid, _ in corners:
        v = valid & ok
        np.add.at(gV, (iz[v], iy[v], ix[v]), (w * gk)[v])
    return gV


def extract_rotation_grad(volume, grad_slices, image_shape, R_zyx, fftfreq_max=None,
                          x_tolerance=1e-8):
    """Gradient of ``Re<grad_slices, extract(volume, R)>`` w.r.t. R_zyx, (n, 3, 3).

    dL/dc_d = Re(conj(g) * kappa(sum_k dw_k/dc_d V[n_k])) with floor/floor+1
    corners; zero where c_d <= 0 or c_d >= N-1 (grid_sample's border clip
    gradient).  c_d = a_d * s * (R_zyx v)_d + b_d, v = (0, fy, fx), so
    dL/dR[d, j] = sum_p dL/dc_{p,d} a_d s_p v_{p,j}.
    """
    vol = np.asarray(volume, dtype=np.complex128)
    g = np.asarray(grad_slices, dtype=np.complex128)
    Di, Hi, Wi = image_shape
    R = np.asarray(R_zyx, dtype=np.float64).reshape(-1, 3, 3)
    g = g.reshape(R.shape[0], Hi, rfft_width(Wi))
    fr = _rotated_frequencies(R, Hi, Wi)
    conj = _extract_conjugate_mask(fr, R, x_tolerance)
    s = np.where(conj, -1.0, 1.0)
    fr = fr * s[..., None]
    c = _to_array_coords(fr, (Di, Hi, Wi))
    mask = fftfreq_mask(Hi, Wi, fftfreq_max)[None]
    shape = vol.shape
    ok = _inside(c, shape) & mask
    cc = np.where(ok[..., None], c, 0.0)
    i0 = np.floor(cc).astype(np.int64)
    t = cc - i0
    dS = np.zeros(c.shape, dtype=np.complex128)  # d(sample)/dc_d before kappa
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                iz, iy, ix = i0[..., 0] + dz, i0[..., 1] + dy, i0[..., 2] + dx
                valid = ok & (iz < shape[0]) & (iy < shape[1]) & (ix < shape[2])
                V = np.where(valid, vol[np.where(valid, iz, 0), np.where(valid, iy, 0),
                                        np.where(valid, ix, 0)], 0.0)
                wz = t[..., 0] if dz else 1 - t[..., 0]
                wy = t[..., 1] if dy else 1 - t[..., 1]
                wx = t[..., 2] if dx else 1 - t[..., 2]
                sz, sy, sx = (1 if dz else -1), (1 if dy else -1), (1 if dx else -1)
                dS[..., 0] += sz * wy * wx * V
                dS[..., 1] += wz * sy * wx * V
                dS[..., 2] += wz * wy * sx * V
    hi = np.array(shape, dtype=np.float64) - 1
    live = ok[..., None] & (c > 0) & (c < hi)
    dSk = np.where(conj[..., None], np.conj(dS), dS)
    dLdc = np.where(live, np.real(np.conj(g)[..., None] * dSk), 0.0)
    a = _coord_scales((Di, Hi, Wi))
    fy, fx = central_slice_frequencies(Hi, Wi)
    vy = np.broadcast_to(fy, (Hi, fx.shape[1]))[None]
    vx = np.broadcast_to(fx, (Hi, fx.shape[1]))[None]
    gR = np.zeros_like(R)
    for d in range(3):
        w = dLdc[..., d] * a[d] * s
        gR[:, d, 1] = np.sum(w * vy, axis=(1, 2))
        gR[:, d, 2] = np.sum(w * vx, axis=(1, 2))
    return gR


def _insert_setup(R_zyx, volume_shape, fftfreq_max):
    D, H, W = volume_shape
    Wr = rfft_width(W)
    R = np.asarray(R_zyx, dtype=np.float64).reshape(-1, 3, 3)
    fr = _rotated_frequencies(R, H, W)
    mask = fftfreq_mask(H, W, fftfreq_max)
    redundant = np.zeros((H, Wr), dtype=bool)
    redundant[: H // 2, 0] = True  # fx == 0 and fy < 0
    mask = mask & ~redundant
    conj = fr[..., 2] < 0
    s = np.where(conj, -1.0, 1.0)
    fr = fr * s[..., None]
    c = _to_array_coords(fr, (D, H, W))
    shape = (D, H, Wr)
    ok = mask[None] & _inside(c, shape)
    return R, conj, s, c, ok, shape


def _insert_corners(c, ok):
    """floor / ceil corners with trilinear weights (TFS/.../_insert_central_slices_rfft_3d.py:298-329)."""
    cc = np.where(ok[..., None], c, 0.0)
    lo = np.floor(cc).astype(np.int64)
    hi = np.ceil(cc).astype(np.int64)
    t = cc - lo
    out = []
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                iz = hi[..., 0] if dz else lo[..., 0]
                iy = hi[..., 1] if dy else lo[..., 1]
                ix = hi[..., 2] if dx else lo[..., 2]
                wz = t[..., 0] if dz else 1 - t[..., 0]
                wy = t[..., 1] if dy else 1 - t[..., 1]
                wx = t[..., 2] if dx else 1 - t[..., 2]
                out.append(((iz, iy, ix), wz * wy * wx, (dz, dy, dx)))
    return out, t, lo, hi


def insert_central_slices(slices, volume_shape, R_zyx, fftfreq_max=None):
    """Trilinear insertion of rfft slices into a 3-D rfft volume (float64).

    TFS/slice_insertion/_insert_central_slices_rfft_3d.py:9-106,213-406:
    skip slice pixels with fx == 0 and fy < 0; conjugate where x' < 0
    (strict); drop points outside [0, N-1]; floor/ceil corners; add w*d to the
    volume and w to the weights; for every corner on the x=0 plane other than
    DC (D//2, H//2, 0) also add w*conj(d) and w at ((D-z)%D, (H-y)%H, 0).
    Returns (volume (D,H,Wr) complex128, weights (D,H,Wr) float64).
    """
    R, conj, s, c, ok, shape = _insert_setup(R_zyx, volume_shape, fftfreq_max)
    D, H, Wr = shape
    d = np.asarray(slices, dtype=np.complex128).reshape(R.shape[0], H, Wr)
    d = np.where(conj, np.conj(d), d)
    V = np.zeros(shape, dtype=np.complex128)
    Wt = np.zeros(shape, dtype=np.float64)
    corners, _, _, _ = _insert_corners(c, ok)
    for (iz, iy, ix), w, _ in corners:
        np.add.at(V, (iz[ok], iy[ok], ix[ok]), (w * d)[ok])
        np.add.at(Wt, (iz[ok], iy[ok], ix[ok]), w[ok])
        mir = ok & (ix == 0) & ~((iz == D // 2) & (iy == H // 2))
        mz, my = (D - iz) % D, (H - iy) % H
        np.add.at(V, (mz[mir], my[mir], ix[mir]), (w * np.conj(d))[mir])
        np.add.at(Wt, (mz[mir], my[mir], ix[mir]), w[mir])
    return V, Wt


def insert_adjoint(grad_volume, volume_shape, R_zyx, fftfreq_max=None):
    """Gradient of ``Re<gV, insert(d, R)[0]>`` w.r.t. the slices d.

    g_{d'} = sum_k w_k gV[n_k] + sum_{k: x=0, not DC} w_k conj(gV[mirror(n_k)]),
    then g_d = conj(g_{d'}) where the point was conjugated; 0 where masked or
    out of bounds.
    """
    R, conj, s, c, ok, shape = _insert_setup(R_zyx, volume_shape, fftfreq_max)
    D, H, Wr = shape
    gV = np.asarray(grad_volume, dtype=np.complex128)
    g = np.zeros(c.shape[:-1], dtype=np.complex128)
    corners, _, _, _ = _insert_corners(c, ok)
    for (iz, iy, ix), w, _ in corners:
        g += np.where(ok, w * gV[iz, iy, ix], 0.0)
        mir = ok & (ix == 0) & ~((iz == D // 2) & (iy == H // 2))
        g += np.where(mir, w * np.conj(gV[(D - iz) % D, (H - iy) % H, ix]), 0.0)
    return np.where(conj, np.conj(g), g)


def insert_rotation_grad(slices, grad_volume, grad_weights, volume_shape, R_zyx, fftfreq_max=None):
    """Gradient of ``Re<gV, V> + <gW, W>`` w.r.t. R_zyx for (V, W) = insert(d, R).

    Torch differentiates the trilinear weights only (corners are integer
    indices).  With floor/ceil corners the two corners along an axis coincide
    when the coordinate is an integer, so the derivative along that axis is 0
    there.
    """
    R, conj, s, c, ok, shape = _insert_setup(R_zyx, volume_shape, fftfreq_max)
    D, H, Wr = shape
    W_ = volume_shape[2]
    d = np.asarray(slices, dtype=np.complex128).reshape(R.shape[0], H, Wr)
    d = np.where(conj, np.conj(d), d)
    gV = np.asarray(grad_volume, dtype=np.complex128)
    gW = np.zeros(shape) if grad_weights is None else np.asarray(grad_weights, dtype=np.float64)
    corners, t, lo, hi = _insert_corners(c, ok)
    dLdc = np.zeros(c.shape, dtype=np.float64)
    for (iz, iy, ix), w, (dz, dy, dx) in corners:
        q = np.real(np.conj(gV[iz, iy, ix]) * d) + gW[iz, iy, ix]
        mir = ok & (ix == 0) & ~((iz == D // 2) & (iy == H // 2))
        mz, my = (D - iz) % D, (H - iy) % H
        q = q + np.where(mir, np.real(np.conj(gV[mz, my, ix]) * np.conj(d)) + gW[mz, my, ix], 0.0)
        q = np.where(ok, q, 0.0)
        wz = t[..., 0] if dz else 1 - t[..., 0]
        wy = t[..., 1] if dy else 1 - t[..., 1]
        wx = t[..., 2] if dx else 1 - t[..., 2]
        sz, sy, sx = (1 if dz else -1), (1 if dy else -1), (1 if dx else -1)
        dLdc[..., 0] += sz * wy * wx * q
        dLdc[..., 1] += wz * sy * wx * q
        dLdc[..., 2] += wz * wy * sx * q
    dLdc = np.where(lo == hi, 0.0, dLdc)
    a = _coord_scales(volume_shape)
    fy, fx = central_slice_frequencies(H, W_)
    vy = np.broadcast_to(fy, (H, Wr))[None]
    vx = np.broadcast_to(fx, (H, Wr))[None]
    gR = np.zeros_like(R)
    for k in range(3):
        w = dLdc[..., k] * a[k] * s
        gR[:, k, 1] = np.sum(w * vy, axis=(1, 2))
        gR[:, k, 2] = np.sum(w * vx, axis=(1, 2))
    return gR


def sample_image_3d(image, coordinates, interpolation="trilinear"):
    """Generic 3-D sampler (TII/image_interpolation_3d.py:11-101).

    image (d,h,w) or (c,d,h,w), real or complex; coordinates (..., 3) zyx in
    array units.  Trilinear uses floor/floor+1 corners; nearest rounds half to
    even (grid_sample's nearbyint).  Points outside [0, N-1] give 0.  Returns
    (...) or (..., c).
    """
    img = np.asarray(image)
    multich = img.ndim == 4
    if img.ndim not in (3, 4):
        raise ValueError("image should have shape (d, h, w) or (c, d, h, w)")
    if not multich:
        img = img[None]
    dt = np.complex128 if np.iscomplexobj(img) else np.float64
    img = img.astype(dt)
    coords = np.asarray(coordinates, dtype=np.float64)
    lead = coords.shape[:-1]
    c = coords.reshape(-1, 3)
    shape = img.shape[1:]
    ok = _inside(c, shape)
    out = np.zeros((c.shape[0], img.shape[0]), dtype=dt)
    if interpolation == "nearest":
        idx = _grid_sample_nearest_index(np.where(ok[:, None], c, 0.0), shape)
        vals = img[:, idx[:, 0], idx[:, 1], idx[:, 2]].T
        out = np.where(ok[:, None], vals, 0.0)
    else:
        for ch in range(img.shape[0]):
            out[:, ch] = _gather_trilinear(img[ch], c, ok)
    out = out.reshape(*lead, img.shape[0])
    return out if multich else out[..., 0]


def _grid_sample_source(c, shape, precision):
    """grid_sample's unclipped source coordinate of array coordinate c
    (TII/grid_sample_utils.py:24 normalises g = c / (0.5 N - 0.5) - 1,
    align_corners=True undoes it as u = ((g + 1) / 2) (N - 1)) in the
    sampler's precision; the round trip can move an integer coordinate by an
    ulp, which decides its cell and one-sided derivative."""
    f = np.dtype(precision).type
    cc = np.asarray(c, dtype=precision)
    N = np.asarray(shape, dtype=precision)
    sc = (f(0.5) * N - f(0.5)).astype(precision)
    with np.errstate(divide="ignore", invalid="ignore"):
        g = ((cc / sc).astype(precision) - f(1.0)).astype(precision)
        u = (((g + f(1.0)).astype(precision) / f(2.0)).astype(precision) * (N - f(1.0))).astype(precision)
    return np.where(N > 1, u, 0).astype(np.float64)


def sample_image_3d_grad(image, coordinates, grad_samples, interpolation="trilinear", precision=np.float32):
    """Gradients of sample_image_3d with respect to the image and the coordinates.

    TII/image_interpolation_3d.py:64-89: grid_sample (border padding,
    align_corners=True) then ``samples[~inside] *= 0``, so that
    * samples outside [0, N-1] pass no gradient at all;
    * trilinear: with u the source coordinate (``_grid_sample_source``, in
      the sampler's ``precision``) clipped to [0, N-1], the image gradient is
      the transposed interpolation, ``gI[corner] += w_corner(u) * g``; the
      coordinate gradient along axis d is
      ``Re(conj(g) * sum_corner dw/du_d * I[corner])``, with dw/du_d the
      derivative of the floor / floor+1 weights, and 0 where the border clip
      holds the coordinate (``u_d <= 0`` or ``u_d >= N-1``: grid_sample's
      clip sets the gradient to 0 there);
    * nearest: the whole gradient goes to the sampled voxel and the
      coordinates get none.
    Multichannel images sum the coordinate gradient over channels; complex
    images follow torch's convention (grad = dL/dRe + i dL/dIm).
    Returns (grad_image with the image's shape, grad_coords (..., 3)).
    """
    img = np.asarray(image)
    multich = img.ndim == 4
    if not multich:
        img = img[None]
    dt = np.complex128 if np.iscomplexobj(img) else np.float64
    img = img.astype(dt)
    coords = np.asarray(coordinates, dtype=np.float64)
    lead = coords.shape[:-1]
    c = coords.reshape(-1, 3)
    g = np.asarray(grad_samples).astype(dt).reshape(c.shape[0], img.shape[0])
    shape = img.shape[1:]
    ok = _inside(c, shape)
    gI = np.zeros_like(img)
    gc = np.zeros_like(c)
    if interpolation == "nearest":
        idx = _grid_sample_nearest_index(np.where(ok[:, None], c, 0.0), shape)
        for ch in range(img.shape[0]):
            np.add.at(gI[ch], (idx[ok, 0], idx[ok, 1], idx[ok, 2]), g[ok, ch])
    else:
        u = _grid_sample_source(np.where(ok[:, None], c, 0.0), shape, precision)
        hi = np.asarray(shape, dtype=np.float64) - 1
        clipped = (u <= 0) | (u >= hi)
        cc = np.clip(u, 0.0, hi)
        i0 = np.floor(cc).astype(np.int64)
        t = cc - i0
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    iz, iy, ix = i0[:, 0] + dz, i0[:, 1] + dy, i0[:, 2] + dx
                    valid = ok & (iz < shape[0]) & (iy < shape[1]) & (ix < shape[2])
                    wz = t[:, 0] if dz else 1 - t[:, 0]
                    wy = t[:, 1] if dy else 1 - t[:, 1]
                    wx = t[:, 2] if dx else 1 - t[:, 2]
                    sz, sy, sx = (1.0 if dz else -1.0), (1.0 if dy else -1.0), (1.0 if dx else -1.0)
                    dw = np.stack([sz * wy * wx, wz * sy * wx, wz * wy * sx], axis=-1)
                    jz, jy, jx = np.where(valid, iz, 0), np.where(valid, iy, 0), np.where(valid, ix, 0)
                    for ch in range(img.shape[0]):
                        v = img[ch][jz, jy, jx]
                        np.add.at(gI[ch], (jz[valid], jy[valid], jx[valid]), (wz * wy * wx * g[:, ch])[valid])
                        gc += np.where(valid[:, None], np.real(np.conj(g[:, ch]) * v)[:, None] * dw, 0.0)
        gc = np.where(clipped, 0.0, gc)
    gc = np.where(ok[:, None], gc, 0.0)
    return (gI if multich else gI[0]), gc.reshape(*lead, 3)


def _grid_sample_nearest_index(c, shape):
    """Nearest index as grid_sample computes it, in float32.

    TII/grid_sample_utils.py:6-26 normalises ``g = c / (0.5 N - 0.5) - 1``;
    grid_sample (align_corners=True, border) unnormalises
    ``u = ((g + 1) / 2) * (N - 1)``, clips to [0, N-1] and rounds half to
    even.  The float32 round trip moves exact .5 ties by an ulp, which decides
    them, so it is reproduced exactly here.
    """
    c32 = np.asarray(c, dtype=np.float32)
    N = np.asarray(shape, dtype=np.float32)
    sc = (np.float32(0.5) * N - np.float32(0.5)).astype(np.float32)
    g = ((c32 / sc).astype(np.float32) - np.float32(1.0)).astype(np.float32)
    u = (((g + np.float32(1.0)).astype(np.float32) / np.float32(2.0)).astype(np.float32)
         * (N - np.float32(1.0))).astype(np.float32)
    u = np.clip(u, np.float32(0.0), N - np.float32(1.0))
    return np.rint(u).astype(np.int64)


def insert_into_image_3d(values, coordinates, image, weights=None, interpolation="trilinear"):
    """Generic 3-D splat, in place (TII/image_interpolation_3d.py:104-275).

    trilinear: floor/ceil corners, image += w*v, weights += w; nearest:
    round half to even, weight 1.  Points outside [0, N-1] are dropped; an
    unknown mode changes nothing.  ``image`` and ``weights`` are float64 /
    complex128 NumPy arrays modified in place and returned.
    """
    img = image
    multich = img.ndim == 4
    vals = np.asarray(values)
    coords = np.asarray(coordinates, dtype=np.float64)
    vshape = vals.shape[:-1] if multich else vals.shape
    if vshape != coords.shape[:-1]:
        raise ValueError("One coordinate triplet is required for each value in data.")
    if coords.shape[-1] != 3:
        raise ValueError("Coordinates must be 3D with shape (..., 3).")
    if weights is None:
        weights = np.zeros(img.shape[-3:], dtype=np.float64)
    img4 = img if multich else img[None]
    v = vals.reshape(-1, img4.shape[0]) if multich else vals.reshape(-1, 1)
    c = coords.reshape(-1, 3)
    ok = _inside(c, img4.shape[1:])
    v, c = v[ok], c[ok]
    if interpolation == "nearest":
        idx = np.rint(c).astype(np.int64)
        for ch in range(img4.shape[0]):
            np.add.at(img4[ch], (idx[:, 0], idx[:, 1], idx[:, 2]), v[:, ch])
        np.add.at(weights, (idx[:, 0], idx[:, 1], idx[:, 2]), 1.0)
    elif interpolation == "trilinear":
        corners, _, _, _ = _insert_corners(c, np.ones(c.shape[0], dtype=bool))
        for (iz, iy, ix), w, _ in corners:
            for ch in range(img4.shape[0]):
                np.add.at(img4[ch], (iz, iy, ix), w * v[:, ch])
            np.add.at(weights, (iz, iy, ix), w)
    return img, weights


def extract_central_slices_multichannel(volume, image_shape, R_zyx, fftfreq_max=None, x_tolerance=1e-8):
    """TFS/slice_extraction/_extract_central_slices_rfft_3d.py:225-367: every
    channel of a (c, d, h, w) volume sampled at the same coordinates, as the
    multichannel grid_sample does; returns (n, c, H_i, Wr_i)."""
    vol = np.asarray(volume)
    return np.stack([extract_central_slices(vol[k], image_shape, R_zyx, fftfreq_max, x_tolerance)
                     for k in range(vol.shape[0])], axis=-3)


def insert_central_slices_multichannel(slices, volume_shape, R_zyx, fftfreq_max=None):
    """TFS/slice_insertion/_insert_central_slices_rfft_3d.py:109-210: (n, c, h, wr)
    slices into (c, d, h, wr) volumes at the same coordinates; one weights
    volume shared by the channels (the weights do not depend on the values)."""
    d = np.asarray(slices)
    vols, wts = [], None
    for k in range(d.shape[-3]):
        v, wts = insert_central_slices(d[..., k, :, :], volume_shape, R_zyx, fftfreq_max)
        vols.append(v)
    return np.stack(vols, axis=0), wts


def project_3d_to_2d_multichannel(volume, R, pad_factor=2.0, fftfreq_max=None, zyx_matrices=False):
    """TFS/project.py:96-177: project_3d_to_2d of every channel with the same
    rotations; (c, d, d, d) -> (..., c, d, d)."""
    vol = np.asarray(volume)
    return np.stack([project_3d_to_2d(vol[k], R, pad_factor, fftfreq_max, zyx_matrices)
                     for k in range(vol.shape[0])], axis=-3)


def _sinc2_grid(shape):
    """sinc(|f|)^2 over a full fftshifted frequency grid (TGU/fftfreq_grid.py:10-81)."""
    axes = [np.fft.fftshift(np.fft.fftfreq(n)) for n in shape]
    grids = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt(sum(g * g for g in grids))
    return np.sinc(r) ** 2


def volume_rfft(volume):
    """fftshift -> rfftn -> fftshift(-3,-2) (TFS/slice_extraction/...:95-98)."""
    v = np.fft.fftshift(np.asarray(volume, dtype=np.float64), axes=(-3, -2, -1))
    v = np.fft.rfftn(v, axes=(-3, -2, -1))
    return np.fft.fftshift(v, axes=(-3, -2))


def project_3d_to_2d(volume, R, pad_factor=2.0, fftfreq_max=None, zyx_matrices=False):
    """TFS/project.py:12-93 in float64: pad, x sinc^2, rfft, extract, irfft, unpad."""
    vol = np.asarray(volume, dtype=np.float64)
    d, h, w = vol.shape[-3:]
    if len({d, h, w}) != 1:
        raise ValueError("all dimensions of volume must be equal.")
    if pad_factor < 1.0:
        raise ValueError("pad_factor must be >= 1.0")
    p = 0
    if pad_factor > 1.0:
        p = int((vol.shape[-1] * (pad_factor - 1.0)) // 2)
        vol = np.pad(vol, p)
    shape = vol.shape
    vol = vol * _sinc2_grid(shape)
    dft = volume_rfft(vol)
    R_in = np.asarray(R)
    Rz = prepare_rotations(R_in, zyx_matrices, 1e-8)
    slices = extract_central_slices(dft, shape, Rz, fftfreq_max, 1e-8)
    slices = np.fft.ifftshift(slices, axes=(-2,))
    proj = np.fft.irfftn(slices, axes=(-2, -1))
    proj = np.fft.ifftshift(proj, axes=(-2, -1))
    if p:
        proj = proj[..., p:-p, p:-p]
    return proj.reshape(*R_in.shape[:-2], *proj.shape[-2:])


def backproject_2d_to_3d(images, R, pad_factor=2.0, fftfreq_max=None, zyx_matrices=False):
    """TFS/backproject.py:11-92 in float64: pad, rfft, insert, /clamp(w,1), irfft, /sinc^2."""
    imgs = np.asarray(images, dtype=np.float64)
    h, w = imgs.shape[-2:]
    if h != w:
        raise ValueError("images must be square.")
    if pad_factor < 1.0:
        raise ValueError("pad_factor must be >= 1.0")
    p = 0
    if pad_factor > 1.0:
        p = int((imgs.shape[-1] * (pad_factor - 1.0)) // 2)
        imgs = np.pad(imgs, [(0, 0)] * (imgs.ndim - 2) + [(p, p), (p, p)])
    h, w = imgs.shape[-2:]
    vshape = (w, w, w)
    f = np.fft.fftshift(imgs, axes=(-2, -1))
    f = np.fft.rfftn(f, axes=(-2, -1))
    f = np.fft.fftshift(f, axes=(-2,))
    Rz = prepare_rotations(R, zyx_matrices, None)
    V, Wt = insert_central_slices(f.reshape(-1, *f.shape[-2:]), vshape, Rz, fftfreq_max)
    V = V / np.maximum(Wt, 1.0)
    V = np.fft.ifftshift(V, axes=(-3, -2))
    V = np.fft.irfftn(V, axes=(-3, -2, -1))
    V = np.fft.ifftshift(V, axes=(-3, -2, -1))
    V = V / _sinc2_grid(V.shape)
    if p:
        V = V[p:-p, p:-p, p:-p]
    return np.real(V)
