"""teamtomo_amd: MI355X (gfx950) Fourier-slice gather/scatter for TeamTomo.

Drop-in replacements for the central-slice operators of
``torch_fourier_slice`` (3-D slices and 2-D lines) and the 3-D / 2-D
sampling and splatting of ``torch_image_interpolation``, running as
hand-written HIP kernels behind a C ABI (``include/teamtomo_amd.h``).  ``install_dropin()`` registers modules
under the reference package names so unchanged callers pick them up.
"""

from __future__ import annotations

import sys
import types

from ._config import deterministic_enabled, set_deterministic, set_gather_path, set_pixel_atomic, set_scatter_path
from .fourier_slice import (
    backproject_2d_to_3d,
    backproject_2d_to_3d_multichannel,
    extract_central_slices_rfft_2d,
    extract_central_slices_rfft_3d,
    extract_central_slices_rfft_3d_multichannel,
    insert_central_slices_rfft_3d,
    insert_central_slices_rfft_3d_multichannel,
    project_2d_to_1d,
    project_3d_to_2d,
    project_3d_to_2d_multichannel,
)
from .batch import SliceBatch
from .fourier_shift import (
    fourier_shift_dft_1d,
    fourier_shift_dft_2d,
    fourier_shift_dft_3d,
    fourier_shift_image_1d,
    fourier_shift_image_2d,
    fourier_shift_image_3d,
)
from .image_interpolation import (
    insert_into_image_1d,
    insert_into_image_2d,
    insert_into_image_3d,
    sample_image_1d,
    sample_image_2d,
    sample_image_3d,
)
from .subpixel_crop import subpixel_crop_2d, subpixel_crop_3d
from .transform_image import (
    affine_transform_image_2d,
    affine_transform_image_3d,
    rotate_then_shift_image_2d,
    shift_then_rotate_image_2d,
)

__all__ = [
    "extract_central_slices_rfft_3d",
    "extract_central_slices_rfft_3d_multichannel",
    "insert_central_slices_rfft_3d",
    "insert_central_slices_rfft_3d_multichannel",
    "project_3d_to_2d",
    "project_3d_to_2d_multichannel",
    "backproject_2d_to_3d",
    "backproject_2d_to_3d_multichannel",
    "extract_central_slices_rfft_2d",
    "project_2d_to_1d",
    "sample_image_3d",
    "insert_into_image_3d",
    "sample_image_2d",
    "insert_into_image_2d",
    "sample_image_1d",
    "insert_into_image_1d",
    "affine_transform_image_3d",
    "affine_transform_image_2d",
    "rotate_then_shift_image_2d",
    "shift_then_rotate_image_2d",
    "subpixel_crop_3d",
    "subpixel_crop_2d",
    "fourier_shift_dft_1d",
    "fourier_shift_dft_2d",
    "fourier_shift_dft_3d",
    "fourier_shift_image_1d",
    "fourier_shift_image_2d",
    "fourier_shift_image_3d",
    "set_deterministic",
    "deterministic_enabled",
    "set_pixel_atomic",
    "set_gather_path",
    "set_scatter_path",
    "install_dropin",
    "SliceBatch",
]

# the names torch_fourier_slice exports (TFS/__init__.py:16-42)
_TFS_NAMES = (
    "extract_central_slices_rfft_3d",
    "extract_central_slices_rfft_3d_multichannel",
    "extract_central_slices_rfft_2d",
    "insert_central_slices_rfft_3d",
    "insert_central_slices_rfft_3d_multichannel",
    "project_3d_to_2d",
    "project_3d_to_2d_multichannel",
    "project_2d_to_1d",
    "backproject_2d_to_3d",
    "backproject_2d_to_3d_multichannel",
)


def install_dropin() -> None:
    """Register ``torch_fourier_slice``, ``torch_image_interpolation``,
    ``torch_transform_image``, ``torch_subpixel_crop`` and ``torch_fourier_shift``
    modules (and the ``slice_extraction`` / ``slice_insertion`` and
    ``fourier_shift_dft`` / ``fourier_shift_image`` submodules) backed by
    this package, so ``from torch_fourier_slice import project_3d_to_2d`` in
    unchanged caller code resolves to the gfx950 implementation."""
    tfs = types.ModuleType("torch_fourier_slice")
    ext = types.ModuleType("torch_fourier_slice.slice_extraction")
    ins = types.ModuleType("torch_fourier_slice.slice_insertion")
    for name in ("extract_central_slices_rfft_3d", "extract_central_slices_rfft_3d_multichannel",
                 "extract_central_slices_rfft_2d"):
        setattr(ext, name, globals()[name])
    for name in ("insert_central_slices_rfft_3d", "insert_central_slices_rfft_3d_multichannel"):
        setattr(ins, name, globals()[name])
    for name in _TFS_NAMES:
        setattr(tfs, name, globals()[name])
    tfs.slice_extraction, tfs.slice_insertion = ext, ins
    tii = types.ModuleType("torch_image_interpolation")
    for name in ("sample_image_3d", "insert_into_image_3d", "sample_image_2d", "insert_into_image_2d",
                 "sample_image_1d", "insert_into_image_1d"):
        setattr(tii, name, globals()[name])
    tti = types.ModuleType("torch_transform_image")
    for name in ("affine_transform_image_3d", "affine_transform_image_2d", "rotate_then_shift_image_2d",
                 "shift_then_rotate_image_2d"):
        setattr(tti, name, globals()[name])
    tsc = types.ModuleType("torch_subpixel_crop")
    tsc.subpixel_crop_3d, tsc.subpixel_crop_2d = subpixel_crop_3d, subpixel_crop_2d
    tfsh = types.ModuleType("torch_fourier_shift")
    fsd = types.ModuleType("torch_fourier_shift.fourier_shift_dft")
    fsi = types.ModuleType("torch_fourier_shift.fourier_shift_image")
    for k in (1, 2, 3):
        for mod, name in ((fsd, f"fourier_shift_dft_{k}d"), (fsi, f"fourier_shift_image_{k}d")):
            setattr(mod, name, globals()[name])
            setattr(tfsh, name, globals()[name])
    tfsh.fourier_shift_dft, tfsh.fourier_shift_image = fsd, fsi
    sys.modules.update({
        "torch_fourier_shift": tfsh,
        "torch_fourier_shift.fourier_shift_dft": fsd,
        "torch_fourier_shift.fourier_shift_image": fsi,
        "torch_subpixel_crop": tsc,
        "torch_transform_image": tti,
        "torch_fourier_slice": tfs,
        "torch_fourier_slice.slice_extraction": ext,
        "torch_fourier_slice.slice_insertion": ins,
        "torch_image_interpolation": tii,
    })
