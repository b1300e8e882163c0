"""FP64 NumPy restatement of torch-subpixel-crop.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, never by the
``teamtomo_amd`` package.

* ``subpixel_crop_3d`` -- packages/primitives/torch-subpixel-crop/src/
  torch_subpixel_crop/subpixel_crop_3d.py:12-108;
* ``subpixel_crop_2d`` -- subpixel_crop_2d.py:12-163;
* ``phase_shift_rfft`` -- torch-fourier-shift's phase ramp on a (not fftshifted)
  rfft, packages/primitives/torch-fourier-shift/src/torch_fourier_shift/
  phase_shift_grids.py:60-127: exp(-2 pi i f . s), f = fftfreq (rfftfreq on the
  last axis).

Patches are cut at round-half-to-even positions (grid_sample nearest at
integer coordinates, zeros outside; the centre is index p//2), multiplied by
the mask, then moved by the remainder round(x) - x: rfftn, the ramp, irfftn
(without an output size, as the reference calls it) or the ramped rfft.
Pinning (tests/test_oracle_crop.py): the golden fixture ``subpixel_crop``
written from the reference, integer positions (a plain crop: no shift), and a
half-pixel shift checked against the shift theorem on a band-limited image.
"""

from __future__ import annotations

import numpy as np


def phase_shift_rfft(dft, shifts, image_shape):
    """dft (..., [d,] h, w//2+1) times exp(-2 pi i f . s), s (..., ndim)."""
    nd = len(image_shape)
    freqs = [np.fft.fftfreq(n) for n in image_shape[:-1]] + [np.fft.rfftfreq(image_shape[-1])]
    grids = np.meshgrid(*freqs, indexing="ij")
    s = np.asarray(shifts, dtype=np.float64)
    ang = sum(grids[k][None] * s.reshape(-1, nd)[:, k].reshape((-1,) + (1,) * nd) for k in range(nd))
    ramp = np.exp(-2j * np.pi * ang)
    return np.asarray(dft).reshape(ramp.shape) * ramp


def _crop(images, ipos, p):
    """images (B, [D,] H, W), ipos (n, B, ndim) int -> (n, B, p...), zeros outside."""
    nd = ipos.shape[-1]
    n, B = ipos.shape[:2]
    off = np.arange(p) - p // 2
    out = np.zeros((n, B) + (p,) * nd)
    for m in range(n):
        for b in range(B):
            idx = [ipos[m, b, k] + off for k in range(nd)]
            inside = np.ones((p,) * nd, dtype=bool)
            for k in range(nd):
                shape = [1] * nd
                shape[k] = p
                inside &= ((idx[k] >= 0) & (idx[k] < images.shape[1 + k])).reshape(shape)
            sub = images[b][np.ix_(*[np.clip(ix, 0, images.shape[1 + k] - 1) for k, ix in enumerate(idx)])]
            out[m, b] = np.where(inside, sub, 0.0)
    return out


def _finish(patches, shifts, p, nd, return_rfft, decenter):
    axes = tuple(range(-nd, 0))
    dft = np.fft.rfftn(patches, axes=axes)
    if return_rfft and decenter:
        shifts = shifts + p / 2
    dft = phase_shift_rfft(dft, shifts, (p,) * nd).reshape(dft.shape)
    if return_rfft:
        return dft
    return np.real(np.fft.irfftn(dft, axes=axes))


def subpixel_crop_3d(image, positions, sidelength, mask=None, return_rfft=False, decenter=False):
    pos = np.asarray(positions, dtype=np.float32).astype(np.float64)
    lead = pos.shape[:-1]
    pos = pos.reshape(-1, 1, 3)
    ipos = np.rint(pos)
    shifts = (ipos - pos).reshape(-1, 3)
    patches = _crop(np.asarray(image, dtype=np.float64)[None], ipos.astype(np.int64), sidelength)[:, 0]
    if mask is not None:
        patches = patches * np.asarray(mask, dtype=np.float64)
    out = _finish(patches, shifts, sidelength, 3, return_rfft, decenter)
    return out.reshape(*lead, *out.shape[-3:])


def subpixel_crop_2d(image, positions, sidelength, mask=None, return_rfft=False, decenter=False):
    img = np.asarray(image, dtype=np.float64)
    batched = img.ndim == 3
    if not batched:
        img = img[None]
    pos = np.asarray(positions, dtype=np.float32).astype(np.float64)
    if not batched:
        pos = pos[..., None, :]
    lead = pos.shape[:-2]
    B = img.shape[0]
    pos = pos.reshape(-1, B, 2)
    ipos = np.rint(pos)
    shifts = (ipos - pos).reshape(-1, 2)
    patches = _crop(img, ipos.astype(np.int64), sidelength)
    if mask is not None:
        patches = patches * np.asarray(mask, dtype=np.float64)
    out = _finish(patches.reshape(-1, sidelength, sidelength), shifts, sidelength, 2, return_rfft, decenter)
    out = out.reshape(*lead, B, *out.shape[-2:])
    return out if batched else out[..., 0, :, :]
