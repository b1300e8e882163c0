"""FP64 NumPy restatement of torch-transform-image's affine warps.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, never by the
``teamtomo_amd`` package.

* ``affine_transform_image_3d`` -- packages/primitives/torch-transform-image/src/
  torch_transform_image/transforms_3d.py:10-40;
* ``affine_transform_image_2d``, ``rotate_then_shift_image_2d``,
  ``shift_then_rotate_image_2d`` -- transforms_2d.py:11-178.

The coordinate grid is np.indices of the output shape (torch-grid-utils
coordinate_grid), homogenised and left-multiplied by the (flipped to zyx / yx
unless told otherwise) float32 matrices, whose batch dimensions broadcast
against the grid as torch.matmul does; the samples come from the oracle's
samplers (``fourier_slice.sample_image_3d``, ``slice_2d.sample_image_2d``).
Pinning (tests/test_oracle_transform.py): the golden fixtures written from the
reference (``transform_image``), the identity matrix (returns the image), an
integer translation (an exact shift with zero fill) and a 90-degree rotation
about the centre (an exact permutation).  ``affine_transform_image_3d_grad``
(the gradients of the 3-D warp through the matmul and the sampler): the
golden torch-autograd gradients (``sample_autograd``) and FP64 central
differences of the matrices.
"""

from __future__ import annotations

import numpy as np

from .fourier_slice import sample_image_3d, sample_image_3d_grad
from .slice_2d import sample_image_2d


def _warp(image, matrices, out_shape, flip, sampler, interpolation):
    k = len(out_shape)
    M = np.asarray(matrices, dtype=np.float32).astype(np.float64)
    if flip:
        M = M.copy()
        M[..., :k, :k] = M[..., :k, :k][..., ::-1, ::-1]
        M[..., :k, k] = M[..., :k, k][..., ::-1]
    grid = np.moveaxis(np.indices(out_shape, dtype=np.float64), 0, -1)  # (*out, k)
    hom = np.concatenate([grid, np.ones(grid.shape[:-1] + (1,))], axis=-1)[..., None]  # (*out, k+1, 1)
    coords = (M @ hom)[..., :k, 0]
    return sampler(image, coords, interpolation)


def affine_transform_image_3d(image, matrices, interpolation, output_shape=None, zyx_matrices=False):
    out_shape = tuple(output_shape) if output_shape else np.asarray(image).shape[-3:]
    return _warp(image, matrices, out_shape, not zyx_matrices, sample_image_3d, interpolation)


def affine_transform_image_3d_grad(image, matrices, grad_out, interpolation, output_shape=None,
                                   zyx_matrices=False, exact_matrices=False):
    """Gradients of affine_transform_image_3d (transforms_3d.py:22-40) with
    respect to the image and the (..., 4, 4) matrices: the sampler's
    gradients at coords = M hom(z, y, x), then dL/dM[i, j] = sum over the
    grid of dL/dc_i hom_j, summed over the grid dimensions the matrices
    broadcast against, and flipped back to xyz unless zyx_matrices.  The
    bottom row gets no gradient (the homogeneous coordinate is dropped).
    ``exact_matrices`` keeps FP64 matrices (for finite differences) instead of
    the reference's float32 ones."""
    out_shape = tuple(output_shape) if output_shape else np.asarray(image).shape[-3:]
    M0 = np.asarray(matrices, dtype=np.float64 if exact_matrices else np.float32).astype(np.float64)
    M = M0.copy()
    if not zyx_matrices:
        M[..., :3, :3] = M[..., :3, :3][..., ::-1, ::-1]
        M[..., :3, 3] = M[..., :3, 3][..., ::-1]
    grid = np.moveaxis(np.indices(out_shape, dtype=np.float64), 0, -1)
    hom = np.concatenate([grid, np.ones(grid.shape[:-1] + (1,))], axis=-1)[..., None]
    coords = (M @ hom)[..., :3, 0]
    gI, gc = sample_image_3d_grad(image, coords, grad_out, interpolation)
    gM = np.zeros(np.broadcast_shapes(M.shape[:-2], out_shape) + (4, 4))
    gM[..., :3, :] = gc[..., :, None] * hom[..., 0][..., None, :]
    # sum over the broadcast dimensions down to the matrices' own shape
    lead = gM.ndim - 2 - (M.ndim - 2)
    gM = gM.sum(axis=tuple(range(lead))) if lead else gM
    for ax, n in enumerate(M.shape[:-2]):
        if n == 1 and gM.shape[ax] != 1:
            gM = gM.sum(axis=ax, keepdims=True)
    if not zyx_matrices:
        gM[..., :3, :3] = gM[..., :3, :3][..., ::-1, ::-1]
        gM[..., :3, 3] = gM[..., :3, 3][..., ::-1]
    return gI, gM


def affine_transform_image_2d(image, matrices, interpolation, output_shape=None, yx_matrices=False):
    out_shape = tuple(output_shape) if output_shape else np.asarray(image).shape[-2:]
    return _warp(image, matrices, out_shape, not yx_matrices, sample_image_2d, interpolation)


def _rot(deg):
    a = np.deg2rad(np.float32(deg)).astype(np.float32)
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], dtype=np.float64)


def _tr(t):
    M = np.eye(3)
    M[:2, 2] = np.broadcast_to(np.asarray(t, dtype=np.float64), (2,))
    return M


def _rotate_shift(image, rotate, shift, rotate_first):
    """transforms_2d.py:97-107 and :165-175: T(c) R T(-s) T(-c) (rotate first)
    or T(c) T(-s) R T(-c), c = (h//2, w//2) when rotating, else 0."""
    h, w = np.asarray(image).shape[-2:]
    c = np.array([h // 2, w // 2], dtype=np.float64) if rotate else 0.0
    s = -np.asarray(shift, dtype=np.float64)
    if rotate_first:
        M = _tr(c) @ _rot(rotate) @ _tr(s) @ _tr(-np.asarray(c))
    else:
        M = _tr(c) @ _tr(s) @ _rot(rotate) @ _tr(-np.asarray(c))
    return M.astype(np.float32)


def rotate_then_shift_image_2d(image, rotate=0, shift=(0, 0), interpolation="bicubic"):
    return affine_transform_image_2d(image, _rotate_shift(image, rotate, shift, True), interpolation,
                                     yx_matrices=True)


def shift_then_rotate_image_2d(image, rotate=0, shift=(0, 0), interpolation="bicubic"):
    return affine_transform_image_2d(image, _rotate_shift(image, rotate, shift, False), interpolation,
                                     yx_matrices=True)
