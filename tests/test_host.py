"""Host-side logic of the drop-in layer that runs without a GPU: argument
validation and errors mirror the reference, and nothing falls back to the CPU."""

import os

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


def test_widened_modules_refuse_cpu_and_mirror_errors():
    """The 2-D analogues, affine warps and subpixel crops: CPU tensors raise,
    and the reference's argument errors come first where it raises them."""
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        tt.extract_central_slices_rfft_2d(torch.zeros(8, 5, dtype=torch.complex64), (8, 8), torch.eye(2)[None])
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        tt.sample_image_2d(torch.zeros(4, 4), torch.zeros(2, 2))
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        tt.affine_transform_image_3d(torch.zeros(4, 4, 4), torch.eye(4), "trilinear")
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        tt.subpixel_crop_3d(torch.zeros(8, 8, 8), torch.zeros(1, 3), 4)
    with pytest.raises(ValueError, match="all dimensions of image must be equal"):
        tt.project_2d_to_1d(torch.zeros(4, 5), torch.eye(2)[None])
    with pytest.raises(ValueError, match="pad_factor"):
        tt.project_2d_to_1d(torch.zeros(4, 4), torch.eye(2)[None], pad_factor=0.5)
    with pytest.raises(ValueError, match="same dtype"):
        tt.insert_into_image_2d(torch.zeros(3, dtype=torch.complex128), torch.zeros(3, 2),
                                torch.zeros(4, 4, dtype=torch.complex64))
    with pytest.raises(ValueError, match="One coordinate pair"):
        tt.insert_into_image_2d(torch.zeros(3), torch.zeros(4, 2), torch.zeros(4, 4))
    with pytest.raises(ValueError, match=r"\(h, w\)"):
        tt.sample_image_2d(torch.zeros(4), torch.zeros(2, 2))


def test_affine_matrix_broadcasting_matches_matmul():
    """The host layout of broadcast matrix batches (transform_image._matrices)
    indexes the same matrix the reference's matmul of the grid would use."""
    from teamtomo_amd.transform_image import _matrices

    g = torch.Generator().manual_seed(0)
    for bshape, grid in (((), (3, 4, 5)), ((2, 1, 1, 1), (3, 4, 5)), ((5,), (3, 4, 5)), ((4, 1), (3, 4, 5)),
                         ((2, 3, 1, 5), (3, 4, 5))):
        M = torch.randn(*bshape, 4, 4, generator=g)
        blk, lead, n, (md, mh, mw) = _matrices(M, "cpu", grid, flip=False)
        full = torch.broadcast_shapes(bshape, grid)
        Mf = M.expand(*full, 4, 4).reshape(-1, *grid, 4, 4)
        assert n == Mf.shape[0] and lead == tuple(full[:-3])
        for t in range(n):
            for z, y, x in ((0, 0, 0), (2, 3, 4), (1, 2, 3)):
                mi = ((t * md + (z if md > 1 else 0)) * mh + (y if mh > 1 else 0)) * mw + (x if mw > 1 else 0)
                assert torch.equal(blk[mi], Mf[t, z, y, x, :3, :])
    with pytest.raises(RuntimeError):
        _matrices(torch.randn(7, 4, 4), "cpu", (3, 4, 5), flip=False)  # does not broadcast
    with pytest.raises(RuntimeError, match="same dtype"):
        _matrices(torch.eye(4, dtype=torch.float64), "cpu", (3, 4, 5), flip=False)


def test_install_dropin_registers_widened_modules():
    import sys

    tt.install_dropin()
    try:
        from torch_fourier_slice import extract_central_slices_rfft_2d, project_2d_to_1d
        from torch_image_interpolation import insert_into_image_1d, insert_into_image_2d, sample_image_1d, sample_image_2d
        from torch_subpixel_crop import subpixel_crop_2d, subpixel_crop_3d
        from torch_transform_image import affine_transform_image_3d, rotate_then_shift_image_2d
        from torch_fourier_shift import fourier_shift_dft_2d, fourier_shift_image_3d
        from torch_fourier_shift.fourier_shift_dft import fourier_shift_dft_3d

        assert fourier_shift_dft_2d is tt.fourier_shift_dft_2d
        assert fourier_shift_dft_3d is tt.fourier_shift_dft_3d
        assert fourier_shift_image_3d is tt.fourier_shift_image_3d
        assert project_2d_to_1d is tt.project_2d_to_1d
        assert extract_central_slices_rfft_2d is tt.extract_central_slices_rfft_2d
        assert sample_image_2d is tt.sample_image_2d and insert_into_image_2d is tt.insert_into_image_2d
        assert sample_image_1d is tt.sample_image_1d and insert_into_image_1d is tt.insert_into_image_1d
        assert affine_transform_image_3d is tt.affine_transform_image_3d
        assert rotate_then_shift_image_2d is tt.rotate_then_shift_image_2d
        assert subpixel_crop_3d is tt.subpixel_crop_3d and subpixel_crop_2d is tt.subpixel_crop_2d
    finally:
        for m in ("torch_fourier_slice", "torch_fourier_slice.slice_extraction",
                  "torch_fourier_slice.slice_insertion", "torch_image_interpolation",
                  "torch_transform_image", "torch_subpixel_crop", "torch_fourier_shift",
                  "torch_fourier_shift.fourier_shift_dft", "torch_fourier_shift.fourier_shift_image"):
            sys.modules.pop(m, None)


def test_fourier_shift_refuses_cpu_and_checks_shapes():
    dft = torch.zeros((4, 3), dtype=torch.complex64)
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        tt.fourier_shift_dft_2d(dft, (4, 4), torch.zeros(2), True, False)
    from teamtomo_amd.fourier_shift import _result_dtype

    c64, c128 = torch.complex64, torch.complex128
    assert _result_dtype(torch.zeros(1, dtype=c64), torch.zeros(1, dtype=torch.int64)) == c64
    assert _result_dtype(torch.zeros(1, dtype=c64), torch.zeros(1, dtype=torch.float64)) == c128
    assert _result_dtype(torch.zeros(1, dtype=c128), torch.zeros(1)) == c128
    assert _result_dtype(torch.zeros(1, dtype=torch.float32), torch.zeros(1)) == c64


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N>1 exits with a message (one GPU per box; sharding is
    designed, not launched), before touching a GPU."""
    import json
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_bench_algorithmic_bytes():
    """The roofline's algorithmic bytes at 256^3 x 4096: 8 B per slice pixel
    + 36 B per rotation + 8 B per volume voxel, for each direction."""
    import importlib.util

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    ext, adj = mod.algorithmic_bytes()
    P, V = 256 * 129, 256 * 256 * 129
    assert ext == adj == 4096 * (8 * P + 36) + 8 * V
    assert abs(ext / 1e9 - 1.150) < 1e-3


def test_bench_traffic_parsing(tmp_path):
    """bench.py --write-traffic reads rocprofv3 counter CSVs: kernel names lose
    'void', the library's namespaces and the argument list; a call's
    dispatches are the last occurrence of its launch sequence."""
    import importlib.util

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    full = "void tt::(anonymous namespace)::brick_kernel<8, 16, 16, 0, false, 1>(float __vector(2) const*, tt::SliceGeom)"
    assert mod._kernel_name(full) == "brick_kernel<8, 16, 16, 0, false, 1>"
    assert mod._kernel_name("tt::(anonymous namespace)::plane_rec_kernel(tt::SliceGeom, long)") == "plane_rec_kernel"
    seq = mod.CALL_KERNELS["extract"]
    hdr = '"Dispatch_Id","Kernel_Name","Counter_Name","Counter_Value"\n'
    lines = []
    did = 0
    for rep in range(3):  # three calls, with a foreign kernel between them
        for k, name in enumerate(seq):
            did += 1
            lines.append(f'{did},"void tt::{name}(int)","FETCH_SIZE",{100 * rep + k}\n')
        did += 1
        lines.append(f'{did},"void at::native::copy(int)","FETCH_SIZE",7\n')
    (tmp_path / "run_counter_collection.csv").write_text(hdr + "".join(lines))
    rows = mod._pmc_dispatches(str(tmp_path), "FETCH_SIZE")
    last = mod._last_call(rows, seq)
    assert [r[1] for r in last] == seq and [r[2] for r in last] == [200.0 + k for k in range(len(seq))]
