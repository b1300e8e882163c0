"""Host-side logic of the drop-in layer that runs without a GPU: argument
validation and errors mirror the reference, and nothing falls back to the CPU."""

import pytest
import torch

import teamtomo_amd as tt


def test_project_backproject_validation_errors():
    with pytest.raises(ValueError, match="all dimensions of volume must be equal"):
        tt.project_3d_to_2d(torch.zeros(4, 4, 5), torch.eye(3)[None])
    with pytest.raises(ValueError, match="pad_factor"):
        tt.project_3d_to_2d(torch.zeros(4, 4, 4), torch.eye(3)[None], pad_factor=0.5)
    with pytest.raises(ValueError, match="images must be square"):
        tt.backproject_2d_to_3d(torch.zeros(1, 4, 5), torch.eye(3)[None])


def test_extract_asserts_ndim():
    with pytest.raises(AssertionError):
        tt.extract_central_slices_rfft_3d(torch.zeros(1, 4, 4, 3, dtype=torch.complex64), (4, 4, 4), torch.eye(3))
    with pytest.raises(AssertionError):
        tt.extract_central_slices_rfft_3d_multichannel(torch.zeros(4, 4, 3, dtype=torch.complex64), (4, 4, 4),
                                                       torch.eye(3))


def test_cpu_tensors_are_refused_not_computed():
    """No CPU fallback: CPU tensors raise instead of silently running elsewhere."""
    V = torch.zeros(8, 8, 5, dtype=torch.complex64)
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        tt.extract_central_slices_rfft_3d(V, (8, 8, 8), torch.eye(3)[None])
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        tt.insert_central_slices_rfft_3d(torch.zeros(1, 8, 5, dtype=torch.complex64), (8, 8, 8), torch.eye(3)[None])
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        tt.sample_image_3d(torch.zeros(4, 4, 4), torch.zeros(2, 3))


def test_insert_into_image_validation_mirrors_reference():
    img = torch.zeros(4, 4, 4, dtype=torch.complex64)
    with pytest.raises(ValueError, match="same dtype"):
        tt.insert_into_image_3d(torch.zeros(3, dtype=torch.complex128), torch.zeros(3, 3), img)
    with pytest.raises(ValueError, match="One coordinate triplet"):
        tt.insert_into_image_3d(torch.zeros(3, dtype=torch.complex64), torch.zeros(4, 3), img)
    with pytest.raises(ValueError, match="3D"):
        tt.insert_into_image_3d(torch.zeros(3, dtype=torch.complex64), torch.zeros(3, 2), img)
    with pytest.raises(ValueError, match=r"\(d, h, w\)"):
        tt.sample_image_3d(torch.zeros(4, 4), torch.zeros(2, 3))


def test_rotation_preparation_is_out_of_place():
    from teamtomo_amd.fourier_slice import _prepare_rotations

    R = torch.tensor([[[1.0, 1e-9, 0.0], [0.0, 1.0, 0.0], [2e-9, 0.0, 1.0]]])
    before = R.clone()
    Rz = _prepare_rotations(R, "cpu", False, 1e-8)
    assert torch.equal(R, before)
    assert Rz[0, 2, 1] == 0.0 and Rz[0, 0, 2] == 0.0  # flipped and clamped
    assert torch.equal(Rz[0], torch.flip(torch.where(R.abs() < 1e-8, 0.0, R)[0], (0, 1)))


def test_install_dropin_registers_reference_names():
    import sys

    tt.install_dropin()
    from torch_fourier_slice import project_3d_to_2d
    from torch_fourier_slice.slice_insertion import insert_central_slices_rfft_3d
    from torch_image_interpolation import sample_image_3d

    assert project_3d_to_2d is tt.project_3d_to_2d
    assert insert_central_slices_rfft_3d is tt.insert_central_slices_rfft_3d
    assert sample_image_3d is tt.sample_image_3d
    for m in ("torch_fourier_slice", "torch_fourier_slice.slice_extraction",
              "torch_fourier_slice.slice_insertion", "torch_image_interpolation"):
        sys.modules.pop(m, None)
