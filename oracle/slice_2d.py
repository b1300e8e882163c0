"""FP64 NumPy restatement of the reference's 2-D analogues of the Fourier-slice path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, never by the
``teamtomo_amd`` package.

Covered (reference roots as in ``oracle/fourier_slice.py``):

* ``extract_central_lines`` -- TFS/slice_extraction/_extract_central_slices_rfft_2d.py:9-92
  with TFS/_grids/_central_line_fftfreq_grid.py:6-31 and TFS/_dft_utils.py:5-72;
* ``extract_central_lines_adjoint`` -- its adjoint w.r.t. the image (what torch
  autograd through grid_sample computes);
* ``project_2d_to_1d`` -- TFS/project.py:180-260;
* ``sample_image_2d`` / ``insert_into_image_2d`` -- TII/image_interpolation_2d.py:11-275;
* ``sample_image_1d`` / ``insert_into_image_1d`` -- TII/image_interpolation_1d.py:11-244
  (grid_sample on the image repeated to height 2 at y = 0, which reduces to
  the 1-D rules written out here).

Conventions: a 2-D rfft image is ``fftshift(rfftn(fftshift(img)), axes=(-2,))``
with DC at ``(H//2, 0)``; ``R_yx`` are (n, 2, 2) rotations after the float32 cast
and the xy->yx flip (``prepare_rotations_2d``); every later step runs in float64
from those float32 values.

Pinning (tests/test_oracle_2d.py): every function against the golden fixtures
written from the reference (tests/golden/make_golden.py, ``extract_2d``,
``sample_image_2d``, ``insert_into_image_2d``); ``extract_central_lines`` also
against the R = I known answer (the line is row H//2 of the image), the
90-degree permutation and the x' < 0 conjugate rule; the adjoint by the FP64
dot-product test; the samplers by exact reproduction of affine functions
(bilinear), of constants by the bicubic sampler and of quadratics by its
weight formula at A = -0.5 (Keys 1981), interpolation at integers, and the
reference tests' known answers (TII tests/test_image_interpolation_2d.py:65-160
and tests/test_image_interpolation_1d.py:66-160); the 1-D sampler also
against the golden-pinned 2-D sampler on the repeated image, the
composition the reference runs.
``fftfreq_max`` of the line extraction: parity unpinned (the reference raises
einops.EinopsError for every value, see ``extract_central_lines``).
"""

from __future__ import annotations

import numpy as np

from .fourier_slice import _snap, rfft_width

__all__ = [
    "prepare_rotations_2d",
    "extract_central_lines",
    "extract_central_lines_adjoint",
    "project_2d_to_1d",
    "sample_image_2d",
    "insert_into_image_2d",
    "sample_image_1d",
    "insert_into_image_1d",
    "image_rfft_2d",
]


def prepare_rotations_2d(R, yx_matrices: bool = False) -> np.ndarray:
    """float32 cast, then the xy->yx flip R_yx = R[::-1, ::-1]
    (TFS/slice_extraction/_extract_central_slices_rfft_2d.py:16,56-57)."""
    R32 = np.array(R, dtype=np.float32, copy=True)
    if not yx_matrices:
        R32 = R32[..., ::-1, ::-1]
    return np.ascontiguousarray(R32)


def image_rfft_2d(image):
    """fftshift -> rfftn -> fftshift(-2) of a real (..., h, w) image (TFS/project.py:234-236)."""
    f = np.fft.fftshift(np.asarray(image, dtype=np.float64), axes=(-2, -1))
    f = np.fft.rfftn(f, axes=(-2, -1))
    return np.fft.fftshift(f, axes=(-2,))


def _line_coords(R_yx, image_shape):
    """Rotated, Friedel-folded array coordinates of every line pixel.

    The line is (y, x) = (0, j/W), j in [0, W//2+1) (the central-line grid,
    TFS/_grids/_central_line_fftfreq_grid.py:6-31, rfft, fftshift a no-op on
    the half axis).  Rotation by R_yx; points with x' < 0 (strict) are negated
    and read conjugated (:61-63, :78-79).  Frequencies become coordinates
    ``(f_y + 0.5) (h-1)/(1-1/h)`` and ``2 f_x (Wr-1)`` (TFS/_dft_utils.py:46-67).
    Returns (n, Wr, 2) coordinates and the (n, Wr) conjugate mask.
    """
    R = np.asarray(R_yx, dtype=np.float64).reshape(-1, 2, 2)
    h, w = image_shape
    wr = rfft_width(w)
    f = np.arange(wr, dtype=np.float64) / w
    fy = R[:, 0, 1, None] * f[None]
    fx = R[:, 1, 1, None] * f[None]
    conj = fx < 0
    s = np.where(conj, -1.0, 1.0)
    c = np.empty(fy.shape + (2,), dtype=np.float64)
    c[..., 0] = (s * fy + 0.5) / (1.0 - 1.0 / h) * (h - 1)
    c[..., 1] = (s * fx * 2) * (wr - 1)
    return _snap(c), conj


def _line_mask(image_shape, fftfreq_max):
    h, w = image_shape
    wr = rfft_width(w)
    if fftfreq_max is None:
        return np.ones(wr, dtype=bool)
    fx = (np.arange(wr).astype(np.float32) * np.float32(1.0 / w)).astype(np.float32)
    return fx <= np.float32(fftfreq_max)


def _bilinear_corners(c, shape):
    """floor / floor+1 corners; corners past N-1 get weight 0 (grid_sample,
    align_corners=True, border padding, for points inside [0, N-1])."""
    i0 = np.floor(c).astype(np.int64)
    t = c - i0
    out = []
    for dy in (0, 1):
        for dx in (0, 1):
            iy, ix = i0[..., 0] + dy, i0[..., 1] + dx
            wgt = (t[..., 0] if dy else 1 - t[..., 0]) * (t[..., 1] if dx else 1 - t[..., 1])
            valid = (iy < shape[0]) & (ix < shape[1])
            out.append((iy, ix, wgt, valid))
    return out


def _inside(c, shape):
    hi = np.array(shape, dtype=np.float64) - 1
    return np.all((c >= 0) & (c <= hi), axis=-1)


def extract_central_lines(image_rfft, image_shape, R_yx, fftfreq_max=None):
    """Central lines of a 2-D fftshifted rfft, one per rotation (complex128).

    TFS/slice_extraction/_extract_central_slices_rfft_2d.py:9-92.  Bilinear
    sampling with zero outside the image array's [0, N-1] box
    (TII/image_interpolation_2d.py:11-96).  ``fftfreq_max``: the reference
    selects with an elementwise (w, 2) mask (:34-36), which flattens the grid
    and raises einops.EinopsError for every value; the evident intent, a
    line masked at |f| <= fftfreq_max (|f| = f_x on the line), is restated
    here (parity unpinned).  Returns (n, W//2+1).
    """
    img = np.asarray(image_rfft, dtype=np.complex128)
    c, conj = _line_coords(R_yx, image_shape)
    ok = _inside(c, img.shape) & _line_mask(image_shape, fftfreq_max)[None]
    cc = np.where(ok[..., None], c, 0.0)
    out = np.zeros(c.shape[:-1], dtype=np.complex128)
    for iy, ix, wgt, valid in _bilinear_corners(cc, img.shape):
        v = valid & ok
        out += np.where(v, wgt, 0.0) * img[np.where(v, iy, 0), np.where(v, ix, 0)]
    out = np.where(conj, np.conj(out), out)
    return np.where(ok, out, 0.0)


def extract_central_lines_adjoint(grad_lines, image_rfft_shape, image_shape, R_yx, fftfreq_max=None):
    """Adjoint of ``extract_central_lines`` w.r.t. the image (complex128):
    each line pixel's (conjugated where it was read conjugated) value splatted
    with its bilinear weights -- the image gradient torch autograd computes
    through grid_sample for the real loss Re<grad, lines>."""
    g = np.asarray(grad_lines, dtype=np.complex128).reshape(-1, rfft_width(image_shape[1]))
    c, conj = _line_coords(R_yx, image_shape)
    ok = _inside(c, image_rfft_shape) & _line_mask(image_shape, fftfreq_max)[None]
    cc = np.where(ok[..., None], c, 0.0)
    gk = np.where(conj, np.conj(g), g)
    out = np.zeros(image_rfft_shape, dtype=np.complex128)
    for iy, ix, wgt, valid in _bilinear_corners(cc, image_rfft_shape):
        v = valid & ok
        np.add.at(out, (iy[v], ix[v]), wgt[v] * gk[v])
    return out


def _sinc2_grid_2d(shape):
    """sinc(|f|)^2 over a full fftshifted 2-D frequency grid (TFS/project.py:218-225)."""
    axes = [np.fft.fftshift(np.fft.fftfreq(n)) for n in shape]
    gy, gx = np.meshgrid(*axes, indexing="ij")
    return np.sinc(np.sqrt(gy * gy + gx * gx)) ** 2


def project_2d_to_1d(image, R, pad_factor=2.0, fftfreq_max=None, yx_matrices=False):
    """TFS/project.py:180-260 in float64: pad, x sinc^2, rfft, central lines,
    irfft, ifftshift, unpad.  Returns (..., d)."""
    img = np.asarray(image, dtype=np.float64)
    h, w = img.shape[-2:]
    if h != w:
        raise ValueError("all dimensions of image must be equal.")
    if pad_factor < 1.0:
        raise ValueError("pad_factor must be >= 1.0")
    p = 0
    if pad_factor > 1.0:
        p = int((img.shape[-1] * (pad_factor - 1.0)) // 2)
        img = np.pad(img, p)
    shape = img.shape
    dft = image_rfft_2d(img * _sinc2_grid_2d(shape))
    R_in = np.asarray(R)
    lines = extract_central_lines(dft, shape, prepare_rotations_2d(R_in, yx_matrices), fftfreq_max)
    proj = np.fft.irfftn(lines, axes=(-1,))
    proj = np.fft.ifftshift(proj, axes=(-1,))
    if p:
        proj = proj[..., p:-p]
    return proj.reshape(*R_in.shape[:-2], proj.shape[-1])


def _grid_sample_unnormalize_f32(c, n):
    """grid_sample's float32 round trip of an array coordinate: the reference
    normalises g = c / (0.5 n - 0.5) - 1 (TII/grid_sample_utils.py:6-26) and
    grid_sample unnormalises ((g + 1) / 2) (n - 1) (align_corners=True)."""
    c32 = np.asarray(c, dtype=np.float32)
    n32 = np.float32(n)
    sc = np.float32(np.float32(0.5) * n32 - np.float32(0.5))
    g = (c32 / sc).astype(np.float32) - np.float32(1.0)
    return ((g.astype(np.float32) + np.float32(1.0)) / np.float32(2.0)).astype(np.float32) * (n32 - np.float32(1.0))


def _cubic_weights(t, A=-0.75):
    """Keys cubic-convolution weights of the taps at -1, 0, 1, 2 for fraction t."""
    def near(x):  # |x| <= 1
        return ((A + 2) * x - (A + 3)) * x * x + 1

    def far(x):  # 1 < |x| < 2
        return ((A * x - 5 * A) * x + 8 * A) * x - 4 * A

    return far(t + 1.0), near(t), near(1.0 - t), far(2.0 - t)


def sample_image_2d(image, coordinates, interpolation="bilinear"):
    """Generic 2-D sampler (TII/image_interpolation_2d.py:11-101).

    image (h, w) or (c, h, w), real or complex; coordinates (..., 2) yx in
    array units.  bilinear: floor/floor+1 corners; nearest: grid_sample's
    float32 round trip, clipped, rounded half to even; bicubic: Keys A=-0.75
    over 4x4 taps, each tap index clipped to [0, N-1] (border padding).
    Points outside [0, N-1] give 0.  Returns (...) or (..., c).
    """
    img = np.asarray(image)
    if img.ndim not in (2, 3):
        raise ValueError("image should have shape (h, w) or (c, h, w)")
    multich = img.ndim == 3
    if not multich:
        img = img[None]
    dt = np.complex128 if np.iscomplexobj(img) else np.float64
    img = img.astype(dt)
    coords = np.asarray(coordinates, dtype=np.float64)
    lead = coords.shape[:-1]
    c = coords.reshape(-1, 2)
    h, w = img.shape[1:]
    ok = _inside(c, (h, w))
    cc = np.where(ok[:, None], c, 0.0)
    out = np.zeros((c.shape[0], img.shape[0]), dtype=dt)
    if interpolation == "nearest":
        iy = np.rint(np.clip(_grid_sample_unnormalize_f32(cc[:, 0], h), 0, h - 1)).astype(np.int64)
        ix = np.rint(np.clip(_grid_sample_unnormalize_f32(cc[:, 1], w), 0, w - 1)).astype(np.int64)
        out = img[:, iy, ix].T
    elif interpolation == "bicubic":
        y0 = np.floor(cc[:, 0]).astype(np.int64)
        x0 = np.floor(cc[:, 1]).astype(np.int64)
        wy = _cubic_weights(cc[:, 0] - y0)
        wx = _cubic_weights(cc[:, 1] - x0)
        for a in range(4):
            yy = np.clip(y0 - 1 + a, 0, h - 1)
            for b in range(4):
                xx = np.clip(x0 - 1 + b, 0, w - 1)
                out += (wy[a] * wx[b])[:, None] * img[:, yy, xx].T
    else:
        for iy, ix, wgt, valid in _bilinear_corners(cc, (h, w)):
            out += np.where(valid, wgt, 0.0)[:, None] * img[:, np.where(valid, iy, 0), np.where(valid, ix, 0)].T
    out = np.where(ok[:, None], out, 0.0).reshape(*lead, img.shape[0])
    return out if multich else out[..., 0]


def insert_into_image_2d(values, coordinates, image, weights=None, interpolation="bilinear"):
    """Generic 2-D splat, in place (TII/image_interpolation_2d.py:104-275).

    bilinear: floor/ceil corners, image += w*v, weights += w; nearest: round
    half to even, weight 1.  Points outside [0, N-1] are dropped; an unknown
    mode changes nothing.  ``image`` / ``weights`` are float64 / complex128
    arrays modified in place and returned.
    """
    img = image
    multich = img.ndim == 3
    vals = np.asarray(values)
    coords = np.asarray(coordinates, dtype=np.float64)
    vshape = vals.shape[:-1] if multich else vals.shape
    if vshape != coords.shape[:-1]:
        raise ValueError("One coordinate pair is required for each value in data.")
    if coords.shape[-1] != 2:
        raise ValueError("Coordinates must be 2D with shape (..., 2).")
    if weights is None:
        weights = np.zeros(img.shape[-2:], dtype=np.float64)
    img3 = img if multich else img[None]
    v = vals.reshape(-1, img3.shape[0]) if multich else vals.reshape(-1, 1)
    c = coords.reshape(-1, 2)
    ok = _inside(c, img3.shape[1:])
    v, c = v[ok], c[ok]
    if interpolation == "nearest":
        iy, ix = np.rint(c[:, 0]).astype(np.int64), np.rint(c[:, 1]).astype(np.int64)
        for ch in range(img3.shape[0]):
            np.add.at(img3[ch], (iy, ix), v[:, ch])
        np.add.at(weights, (iy, ix), 1.0)
    elif interpolation == "bilinear":
        y0, x0 = np.floor(c[:, 0]).astype(np.int64), np.floor(c[:, 1]).astype(np.int64)
        y1, x1 = np.ceil(c[:, 0]).astype(np.int64), np.ceil(c[:, 1]).astype(np.int64)
        ty, tx = c[:, 0] - y0, c[:, 1] - x0
        for yy, xx, wgt in ((y0, x0, (1 - ty) * (1 - tx)), (y0, x1, (1 - ty) * tx),
                            (y1, x0, ty * (1 - tx)), (y1, x1, ty * tx)):
            for ch in range(img3.shape[0]):
                np.add.at(img3[ch], (yy, xx), wgt * v[:, ch])
            np.add.at(weights, (yy, xx), wgt)
    return img, weights


def sample_image_1d(image, coordinates, interpolation="linear"):
    """Generic 1-D sampler (TII/image_interpolation_1d.py:11-107).

    image (w,) or (c, w), real or complex; coordinates (...) in array units.
    The reference samples the image repeated to height 2 at y = 0 with border
    padding, so every y tap reads the same row: linear is (1 - t) v[x0] +
    t v[x0 + 1]; nearest rounds grid_sample's float32 round trip half to
    even; cubic is Keys A=-0.75 over the taps x0-1..x0+2 clipped to [0, w-1].
    Points outside [0, w-1] give 0.  ``interpolation`` is 'linear' (or
    'bilinear'), 'cubic' (or 'bicubic') or 'nearest'.  Returns (...) or (..., c).
    """
    img = np.asarray(image)
    if img.ndim not in (1, 2):
        raise ValueError("image should have shape (w, ) or (c, w)")
    multich = img.ndim == 2
    if not multich:
        img = img[None]
    dt = np.complex128 if np.iscomplexobj(img) else np.float64
    img = img.astype(dt)
    coords = np.asarray(coordinates, dtype=np.float64)
    lead = coords.shape
    c = coords.reshape(-1)
    w = img.shape[1]
    ok = (c >= 0) & (c <= w - 1)
    cc = np.where(ok, c, 0.0)
    mode = {"bilinear": "linear", "bicubic": "cubic"}.get(interpolation, interpolation)
    if mode == "nearest":
        ix = np.rint(np.clip(_grid_sample_unnormalize_f32(cc, w), 0, w - 1)).astype(np.int64)
        out = img[:, ix].T
    elif mode == "cubic":
        x0 = np.floor(cc).astype(np.int64)
        wx = _cubic_weights(cc - x0)
        out = np.zeros((c.shape[0], img.shape[0]), dtype=dt)
        for b in range(4):
            out += wx[b][:, None] * img[:, np.clip(x0 - 1 + b, 0, w - 1)].T
    elif mode == "linear":
        x0 = np.floor(cc).astype(np.int64)
        x1 = np.minimum(x0 + 1, w - 1)
        t = (cc - x0)[:, None]
        out = (1.0 - t) * img[:, x0].T + t * img[:, x1].T
    else:
        raise ValueError(f"unknown interpolation {interpolation!r}")
    out = np.where(ok[:, None], out, 0.0).reshape(*lead, img.shape[0])
    return out if multich else out[..., 0]


def insert_into_image_1d(values, coordinates, image, weights=None, interpolation="linear"):
    """Generic 1-D splat, in place (TII/image_interpolation_1d.py:110-244).

    linear: floor/ceil corners with weights 1 - t and t, image += w*v,
    weights += w; nearest: round half to even, weight 1.  Points outside
    [0, w-1] are dropped; an unknown mode changes nothing.  ``image`` (w,) or
    (c, w) and ``weights`` (w,) are float64 / complex128 arrays modified in
    place and returned.
    """
    img = image
    multich = img.ndim == 2
    vals = np.asarray(values)
    coords = np.asarray(coordinates, dtype=np.float64)
    vshape = vals.shape[:-1] if multich else vals.shape
    if vshape != coords.shape:
        raise ValueError("One coordinate is required for each value in data.")
    w = img.shape[-1]
    if weights is None:
        weights = np.zeros((w,), dtype=np.float64)
    img2 = img if multich else img[None]
    v = vals.reshape(-1, img2.shape[0]) if multich else vals.reshape(-1, 1)
    c = coords.reshape(-1)
    ok = (c >= 0) & (c <= w - 1)
    v, c = v[ok], c[ok]
    if interpolation == "nearest":
        ix = np.rint(c).astype(np.int64)
        for ch in range(img2.shape[0]):
            np.add.at(img2[ch], ix, v[:, ch])
        np.add.at(weights, ix, 1.0)
    elif interpolation == "linear":
        x0, x1 = np.floor(c).astype(np.int64), np.ceil(c).astype(np.int64)
        t = c - x0
        for xx, wgt in ((x0, 1 - t), (x1, t)):
            for ch in range(img2.shape[0]):
                np.add.at(img2[ch], xx, wgt * v[:, ch])
            np.add.at(weights, xx, wgt)
    return img, weights
