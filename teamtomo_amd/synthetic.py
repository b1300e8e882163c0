"""Seeded synthetic inputs for benchmarks and tests (no data files needed).

* ``uniform_rotations``: q = normalize(N(0, I4)) -> 3x3 xyz matrix (uniform on SO(3)).
* ``blob_volume``: a sum of Gaussian blobs plus white noise.
* ``volume_rfft``: the reference's preprocessing, fftshift -> rfftn -> fftshift(-3,-2)
  (slice_extraction/_extract_central_slices_rfft_3d.py:95-98).
"""

from __future__ import annotations

import torch


def uniform_rotations(n: int, seed: int, device="cpu", dtype=torch.float32) -> torch.Tensor:
    g = torch.Generator().manual_seed(seed)
    q = torch.randn(n, 4, generator=g, dtype=torch.float64)
    q = q / q.norm(dim=-1, keepdim=True)
    w, x, y, z = q.unbind(-1)
    R = torch.stack(
        [
            1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
            2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
            2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y),
        ],
        dim=-1,
    ).reshape(n, 3, 3)
    return R.to(device=device, dtype=dtype)


def blob_volume(n: int, n_blobs: int, seed: int, sigma=(2.0, 8.0), amp=(0.5, 1.0),
                noise: float = 0.05, device="cpu") -> torch.Tensor:
    """float32 (n, n, n): n_blobs Gaussians (centres uniform in the central 60%
    of the box, sigma ~ U(sigma), amplitude ~ U(amp)) plus N(0, noise^2)."""
    g = torch.Generator().manual_seed(seed)
    centres = (0.2 + 0.6 * torch.rand(n_blobs, 3, generator=g, dtype=torch.float64)) * n
    sig = sigma[0] + (sigma[1] - sigma[0]) * torch.rand(n_blobs, generator=g, dtype=torch.float64)
    a = amp[0] + (amp[1] - amp[0]) * torch.rand(n_blobs, generator=g, dtype=torch.float64)
    noise_t = noise * torch.randn(n, n, n, generator=g, dtype=torch.float32)
    ax = torch.arange(n, dtype=torch.float32, device=device)
    vol = noise_t.to(device)
    for k in range(n_blobs):
        cz, cy, cx = (float(v) for v in centres[k])
        s2 = 2.0 * float(sig[k]) ** 2
        gz = torch.exp(-((ax - cz) ** 2) / s2)
        gy = torch.exp(-((ax - cy) ** 2) / s2)
        gx = torch.exp(-((ax - cx) ** 2) / s2)
        vol += float(a[k]) * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return vol


def volume_rfft(volume: torch.Tensor) -> torch.Tensor:
    f = torch.fft.fftshift(volume, dim=(-3, -2, -1))
    f = torch.fft.rfftn(f, dim=(-3, -2, -1))
    return torch.fft.fftshift(f, dim=(-3, -2))
