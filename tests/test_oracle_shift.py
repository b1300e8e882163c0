"""Pins the FP64 Fourier-shift oracle (oracle/fourier_shift.py).

Known answers: the phase grids and shifted images of the reference's tests
(packages/primitives/torch-fourier-shift/tests/test_phase_shift_1d.py,
_2d.py, _3d.py; their float32 values are compared with the same allclose
tolerance the reference uses).  Mathematics: integer shifts equal np.roll
(the shift theorem; a sign or axis mistake rolls the wrong way), the
fftshifted layout shifts the same image once the layout is undone, the rfft
grid is the first w//2+1 columns of the full grid, a subpixel shift of an
odd-length signal equals Dirichlet-kernel interpolation, and the shift
gradient matches central finite differences.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fourier_shift as O


def _allclose(a, b, rtol=1e-5, atol=1e-8):
    return np.allclose(np.asarray(a), np.asarray(b), rtol=rtol, atol=atol)


# values from tests/test_phase_shift_{1,2,3}d.py (float32 torch results)
REF_1D = np.array([[1.0 - 0.0j, -1.0 - 8.7423e-08j], [1.0 - 0.0j, 1.0 + 1.7485e-07j]])
REF_2D = np.array([[[1 + 0.0j, 1 + 1.7485e-07j], [-1 - 8.7423e-08j, -1 - 2.3850e-08j]]])
REF_3D_FULL = np.array([[[1.0 + 0.0j, -1.0 - 2.3850e-08j], [1.0 + 1.7485e-07j, -1.0 - 6.7553e-07j]],
                        [[-1.0 - 8.7423e-08j, 1.0 + 3.4969e-07j], [-1.0 - 2.3850e-08j, 1.0 + 4.7700e-08j]]])
REF_3D_RFFT = np.array([[[1.0 + 0.0j, -1.0 + 2.3850e-08j], [1.0 + 1.7485e-07j, -1.0 - 1.5100e-07j]],
                        [[-1.0 - 8.7423e-08j, 1.0 + 3.0199e-07j], [-1.0 - 2.3850e-08j, 1.0 + 0.0j]]])


def test_reference_phase_grids():
    assert _allclose(O.phase_shift_grid(np.zeros((1, 2)), (2,), rfft=False), np.ones((1, 2, 2)))
    assert O.phase_shift_grid(np.zeros((1, 2)), (2,), rfft=True).shape == (1, 2, 2)
    assert _allclose(O.phase_shift_grid(np.array([1, 2]), (2,), rfft=False), REF_1D)
    assert _allclose(O.phase_shift_grid(np.zeros((1, 2)), (2, 2), rfft=False), np.ones((1, 2, 2)))
    assert O.phase_shift_grid(np.zeros((1, 2)), (2, 2), rfft=True).shape == (1, 2, 2)
    assert _allclose(O.phase_shift_grid(np.array([[1, 2]]), (2, 2), rfft=False), REF_2D)
    assert O.phase_shift_grid(np.zeros((1, 3)), (2, 2, 2), rfft=True).shape == (1, 2, 2, 2)
    assert _allclose(O.phase_shift_grid(np.array([[1, 2, 3]]), (2, 2, 2), rfft=False)[0], REF_3D_FULL)
    assert _allclose(O.phase_shift_grid(np.array([[1, 2, 3]]), (2, 2, 2), rfft=True)[0], REF_3D_RFFT)


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_reference_delta_shifts(nd):
    """A delta at index 2 of a 4^nd image moves by +1 along each shifted axis
    (and by -1 in 1-D), as the reference's image tests expect."""
    img = np.zeros((4,) * nd)
    img[(2,) * nd] = 1.0
    cases = [np.ones(nd)] + [np.eye(nd)[a] for a in range(nd)]
    if nd == 1:
        cases = [np.ones(1), -np.ones(1)]
    for s in cases:
        shifts = s[None] if nd > 1 else s
        out = O.fourier_shift_image_nd(img, shifts, nd)[0]
        want = np.zeros_like(img)
        want[tuple(int(2 + v) for v in s)] = 1.0
        assert np.allclose(out, want, atol=1e-5), (nd, s)


@pytest.mark.parametrize("shape", [(12,), (8, 10), (6, 8, 10)])
def test_integer_shift_is_roll(shape):
    nd = len(shape)
    rng = np.random.default_rng(nd)
    img = rng.standard_normal((3,) + shape)
    s = rng.integers(-7, 8, size=(3, nd)).astype(np.float64)
    out = O.fourier_shift_image_nd(img, s if nd > 1 else s[:, 0], nd)
    for b in range(3):
        want = np.roll(img[b], tuple(int(v) for v in s[b]), axis=tuple(range(nd)))
        np.testing.assert_allclose(out[b], want, atol=1e-12)


@pytest.mark.parametrize("rfft", [False, True])
def test_fftshifted_layout_shifts_the_same_image(rfft):
    rng = np.random.default_rng(5)
    shape = (6, 8)
    img = rng.standard_normal(shape)
    s = np.array([[2.0, -3.0]])
    axes = (-2, -1) if not rfft else (-2,)
    dft = np.fft.rfft2(img) if rfft else np.fft.fft2(img)
    shifted = O.fourier_shift_dft(np.fft.fftshift(dft, axes=axes), shape, s, rfft, True)[0]
    back = np.fft.ifftshift(shifted, axes=axes)
    out = np.fft.irfft2(back, s=shape) if rfft else np.real(np.fft.ifft2(back))
    np.testing.assert_allclose(out, np.roll(img, (2, -3), axis=(0, 1)), atol=1e-12)


def test_rfft_grid_is_half_of_full_grid():
    """Odd w: rfftfreq(w) is the first w//2+1 values of fftfreq(w).  Even w:
    the same but for the Nyquist column, where rfftfreq has +1/2 and fftfreq
    -1/2 (conjugate phases)."""
    s = np.array([[0.3, -1.7, 2.2]])
    full = O.phase_shift_grid(s, (5, 6, 9), rfft=False)
    half = O.phase_shift_grid(s, (5, 6, 9), rfft=True)
    np.testing.assert_allclose(half, full[..., :5], atol=1e-14)
    full = O.phase_shift_grid(s, (5, 6, 8), rfft=False)
    half = O.phase_shift_grid(s, (5, 6, 8), rfft=True)
    np.testing.assert_allclose(half[..., :4], full[..., :4], atol=1e-14)
    ny = np.exp(-2j * np.pi * (0.5 * 2.2)) * O.phase_shift_grid(s[:, :2], (5, 6), rfft=False)
    np.testing.assert_allclose(half[..., 4], ny, atol=1e-14)


def test_subpixel_shift_is_dirichlet_interpolation():
    """Odd N: the full-DFT shift by s is x convolved with the periodic sinc
    D(t) = sin(pi t) / (N sin(pi t / N)) at t = n - s - m."""
    n = 11
    rng = np.random.default_rng(9)
    x = rng.standard_normal(n)
    s = 0.37
    y = np.fft.ifft(O.fourier_shift_dft(np.fft.fft(x), (n,), np.array(s), False, False))
    t = np.arange(n)[:, None] - s - np.arange(n)[None, :]
    D = np.sin(np.pi * t) / (n * np.sin(np.pi * t / n))
    np.testing.assert_allclose(np.real(y), D @ x, atol=1e-12)
    assert np.max(np.abs(np.imag(y))) < 1e-12


@pytest.mark.parametrize("rfft,fftshifted", [(False, False), (True, False), (False, True), (True, True)])
def test_shift_grad_matches_finite_differences(rfft, fftshifted):
    rng = np.random.default_rng(11)
    shape = (5, 6)
    stored = (5, 4) if rfft else shape
    dft = rng.standard_normal((2,) + stored) + 1j * rng.standard_normal((2,) + stored)
    g = rng.standard_normal((2,) + stored) + 1j * rng.standard_normal((2,) + stored)
    s = rng.uniform(-3, 3, size=(2, 2))

    def loss(sh):
        return float(np.sum(np.real(np.conj(g) * O.fourier_shift_dft(dft, shape, sh, rfft, fftshifted))))

    got = O.shift_grad(g, dft, s, shape, rfft, fftshifted)
    h = 1e-6
    for b in range(2):
        for a in range(2):
            e = np.zeros_like(s)
            e[b, a] = h
            fd = (loss(s + e) - loss(s - e)) / (2 * h)
            assert abs(got[b, a] - fd) < 1e-6 * max(1.0, abs(fd)), (b, a, got[b, a], fd)
