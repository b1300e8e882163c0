"""ctypes wrapper of the FP64 C oracle (oracle/c/tt_oracle.c).

TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py's CPU baseline).  Same
semantics as oracle.fourier_slice; built by ``make -C oracle/c``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libtt_oracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "c")], check=True)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        p, i, l, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        lib.tt_oracle_extract.argtypes = [p, i, i, i, i, i, i, p, l, d, d, p]
        lib.tt_oracle_extract_adjoint.argtypes = [p, i, i, i, i, i, i, p, l, d, d, p]
        lib.tt_oracle_insert.argtypes = [p, i, i, i, p, l, d, p, p]
        lib.tt_oracle_threads.restype = i
        _lib = lib
    return _lib


def threads() -> int:
    return _load().tt_oracle_threads()


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _rot(R_zyx):
    return np.ascontiguousarray(np.asarray(R_zyx, dtype=np.float32).reshape(-1, 3, 3))


def _fm(fftfreq_max):
    return -1.0 if fftfreq_max is None else float(fftfreq_max)


def extract(volume, image_shape, R_zyx, fftfreq_max=None, x_tolerance=1e-8):
    V = _c128(volume)
    R = _rot(R_zyx)
    Hi, Wri = image_shape[1], int(image_shape[2] / 2 + 1)
    out = np.empty((R.shape[0], Hi, Wri), dtype=np.complex128)
    _load().tt_oracle_extract(V.ctypes.data, *V.shape, *image_shape, R.ctypes.data, R.shape[0],
                              _fm(fftfreq_max), float(x_tolerance), out.ctypes.data)
    return out


def extract_adjoint(grad_slices, volume_shape_rfft, image_shape, R_zyx, fftfreq_max=None, x_tolerance=1e-8):
    g = _c128(grad_slices)
    R = _rot(R_zyx)
    gV = np.zeros(volume_shape_rfft, dtype=np.complex128)
    _load().tt_oracle_extract_adjoint(g.ctypes.data, *volume_shape_rfft, *image_shape, R.ctypes.data,
                                      R.shape[0], _fm(fftfreq_max), float(x_tolerance), gV.ctypes.data)
    return gV


def insert(slices, volume_shape, R_zyx, fftfreq_max=None):
    d = _c128(slices)
    R = _rot(R_zyx)
    D, H, W = volume_shape
    V = np.zeros((D, H, int(W / 2 + 1)), dtype=np.complex128)
    Wt = np.zeros(V.shape, dtype=np.float64)
    _load().tt_oracle_insert(d.ctypes.data, D, H, W, R.ctypes.data, R.shape[0], _fm(fftfreq_max),
                             V.ctypes.data, Wt.ctypes.data)
    return V, Wt
