"""FP64 NumPy restatement of torch-fourier-shift.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, never by the
``teamtomo_amd`` package.

* ``phase_shift_grid`` -- packages/primitives/torch-fourier-shift/src/
  torch_fourier_shift/phase_shift_grids.py:9-127: angle = -2 pi sum_a f_a s_a
  with f = fftfreq on every axis (rfftfreq on the last for an rfft,
  fftfreq_grids.py:9-116), exp(i angle), then fftshift over every axis, or
  every axis but the last for an rfft (dft_utils.py:12-54);
* ``fourier_shift_dft`` -- fourier_shift_dft.py:10-115: the DFT times that
  grid, batch dimensions broadcast;
* ``fourier_shift_image_nd`` -- fourier_shift_image.py:10-91: rfftn, the
  product with rfft=True, fftshifted=False, irfftn without an output size;
* ``shift_grad`` -- the gradient of that product with respect to the shifts
  (torch autograd through the reference's cos/sin grid).

Pinning (tests/test_oracle_shift.py): the phase grids and image shifts the
reference's own tests hold (tests/test_phase_shift_1d.py, _2d.py, _3d.py),
an integer shift equal to np.roll of the image (the shift theorem, which a
sign or axis mistake breaks), the fftshifted layout shifting the same image
once the layout is undone, the rfft grid against the full grid (equal but
for the Nyquist column of an even w), Dirichlet-kernel interpolation of an
odd-length signal, and the shift gradient against finite differences.
"""

from __future__ import annotations

import numpy as np


def _shift_axes(ndim, rfft):
    return tuple(range(-ndim, 0)) if not rfft else tuple(range(-ndim, -1))


def phase_shift_grid(shifts, image_shape, rfft=False, fftshift=False):
    """(..., *stored) complex128 phase grid for (..., ndim) shifts ((...) in 1-D)."""
    nd = len(image_shape)
    s = np.asarray(shifts, dtype=np.float64)
    if nd == 1:
        s = s[..., None]
    freqs = [np.fft.fftfreq(n) for n in image_shape[:-1]]
    freqs.append(np.fft.rfftfreq(image_shape[-1]) if rfft else np.fft.fftfreq(image_shape[-1]))
    grids = np.meshgrid(*freqs, indexing="ij")
    lead = s.shape[:-1]
    ang = np.zeros(lead + grids[0].shape)
    for a in range(nd):
        ang = ang + grids[a] * s[..., a].reshape(lead + (1,) * nd)
    ramp = np.exp(-2j * np.pi * ang)
    if fftshift:
        ramp = np.fft.fftshift(ramp, axes=_shift_axes(nd, rfft))
    return ramp


def fourier_shift_dft(dft, image_shape, shifts, rfft, fftshifted):
    """dft (..., *stored) times the phase grid of the shifts, broadcast."""
    return np.asarray(dft) * phase_shift_grid(shifts, image_shape, rfft, fftshifted)


def _shift_image(image, s, nd):
    axes = tuple(range(-nd, 0))
    dft = np.fft.rfftn(image, axes=axes)
    dft = fourier_shift_dft(dft, image.shape[-nd:], s, True, False)
    return np.fft.irfftn(dft, axes=axes)


def fourier_shift_image_nd(image, shifts, nd):
    """As ``fourier_shift_image`` with the dimensionality given."""
    return _shift_image(np.asarray(image, dtype=np.float64), np.asarray(shifts, dtype=np.float64), nd)


def shift_grad(grad_out, dft, shifts, image_shape, rfft, fftshifted):
    """Gradient of sum Re(conj(grad_out) * out) with respect to the shifts
    (real, (..., ndim)) for out = dft * phase grid: d out / d s_a =
    -2 pi i f_a out, so grad_s_a = 2 pi sum_v f_a(v) Im(conj(grad_out) out)."""
    nd = len(image_shape)
    out = fourier_shift_dft(dft, image_shape, shifts, rfft, fftshifted)
    c = np.imag(np.conj(grad_out) * out)
    freqs = [np.fft.fftfreq(n) for n in image_shape[:-1]]
    freqs.append(np.fft.rfftfreq(image_shape[-1]) if rfft else np.fft.fftfreq(image_shape[-1]))
    grids = np.meshgrid(*freqs, indexing="ij")
    if fftshifted:
        grids = [np.fft.fftshift(g, axes=_shift_axes(nd, rfft)) for g in grids]
    axes = tuple(range(-nd, 0))
    return np.stack([2 * np.pi * np.sum(c * g, axis=axes) for g in grids], axis=-1)
