"""The allocation-free extraction (tt_extract_c64_noalloc) against
tt_extract_c64 through the C ABI: bitwise equal slices on two calls in a row,
the second smaller than the first, and a clear error beyond its fixed
scratch."""

import pytest
import torch

from teamtomo_amd import _lib
from teamtomo_amd.fourier_slice import _prepare_rotations
from teamtomo_amd.synthetic import blob_volume, uniform_rotations, volume_rfft

pytestmark = pytest.mark.gpu


def _extract(fn, V, n, Rz, fmax, out):
    _lib.call(fn, _lib.ptr(V), V.shape[0], V.shape[1], V.shape[2], n, n, n, _lib.ptr(Rz), Rz.shape[0],
              fmax, 1e-8, _lib.ptr(out), _lib.stream_of(V))


def test_noalloc_extract_equals_tt_extract_c64(device):
    cases = [(128, 1500, -1.0, 20), (64, 300, 0.35, 21)]  # the second call is smaller
    for n, n_rot, fmax, seed in cases:
        print(f"noalloc extraction: {n}^3, {n_rot} rotations, fftfreq_max {fmax}", flush=True)
        V = volume_rfft(blob_volume(n, 40, seed=seed, device=device)).contiguous()
        R = uniform_rotations(n_rot, seed + 100, device=device)
        Rz = _prepare_rotations(R, device, False, 1e-8).reshape(-1, 3, 3).contiguous()
        ref = torch.empty((n_rot, n, n // 2 + 1), dtype=torch.complex64, device=device)
        got = torch.full_like(ref, float("nan"))
        _extract("tt_extract_c64", V, n, Rz, fmax, ref)
        _extract("tt_extract_c64_noalloc", V, n, Rz, fmax, got)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), f"{n}^3 x {n_rot}: noalloc slices differ"


def test_noalloc_refuses_beyond_its_scratch(device):
    n = 32
    V = torch.zeros((n, n, n // 2 + 1), dtype=torch.complex64, device=device)
    Rz = torch.zeros((16385, 3, 3), dtype=torch.float32, device=device)
    out = torch.empty((16385, n, n // 2 + 1), dtype=torch.complex64, device=device)
    with pytest.raises(RuntimeError, match="at most 16384 rotations"):
        _extract("tt_extract_c64_noalloc", V, n, Rz, -1.0, out)
