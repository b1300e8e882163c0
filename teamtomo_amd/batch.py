"""Batched central-slice extraction and insertion over many volumes.

A loop of ``extract_central_slices_rfft_3d`` or ``insert_central_slices_rfft_3d``
calls, one per volume (each with its own box size and rotation count), run as
one persistent work-queue launch of the brick kernel per tile configuration
(``tt_batch_*`` in include/teamtomo_amd.h).  The batch is planned and its
scratch allocated once; ``run()`` only launches kernels on the current stream,
so it can be captured in a ``torch.cuda.CUDAGraph`` and replayed.  The input,
rotation and output tensors are bound at construction: update the inputs or
the rotations in place and run (or replay) again.  Every run re-reads the
caller's rotation tensors (the cast, near-zero clamp and zyx flip run as the
first kernel of the run), so inputs and rotations must be contiguous tensors
on the batch's device, rotations float32 or float64 of shape (n, 3, 3).

The tile shape (brick of cells) of each box size is chosen by timing the
candidates on the batch's own items of that size (``autotune=True``);
``tile_shapes`` reports the choice.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .fourier_slice import _fmax_arg, _prepare_rotations, _rfft_width


class SliceBatch:
    """Batched extraction (``kind="extract"``) or insertion (``kind="insert"``).

    extract: ``inputs`` are fftshifted rfft volumes ``(D, H, W//2+1)`` complex64,
    ``shapes`` the image shapes (default: each volume's full shape); ``outputs``
    are the slices ``(n, H, W//2+1)``.
    insert: ``inputs`` are slices ``(n, H, W//2+1)`` complex64, ``shapes`` the
    volume shapes ``(D, H, W)``; ``outputs`` are ``(volume, weights)`` pairs.
    ``rotations`` are ``(n, 3, 3)`` matrices per item, as for the one-call
    functions (xyz unless ``zyx_matrices``).
    """

    def __init__(self, kind, inputs, rotations, shapes=None, fftfreq_max=None, zyx_matrices=False,
                 x_tolerance=1e-8, rot_tolerance=1e-8, autotune=True):
        if kind not in ("extract", "insert"):
            raise ValueError("kind must be 'extract' or 'insert'")
        if len(inputs) != len(rotations) or not inputs:
            raise ValueError("one rotation stack per input is required")
        self.kind = kind
        self._handle = None
        self._keep = []
        self.outputs = []
        items = (_lib.BatchItem * len(inputs))()
        self._rot = []  # (caller's rotations, prepared buffer) per item, re-prepared by every run
        self._flip = int(not zyx_matrices)
        self._rot_tol = -1.0 if (kind != "extract" or rot_tolerance is None) else float(rot_tolerance)
        for k, (x, R) in enumerate(zip(inputs, rotations)):
            _lib.require_gpu(x)
            if x.dtype != torch.complex64:
                raise TypeError(f"SliceBatch: complex64 input required, got {x.dtype}")
            if not x.is_contiguous():
                raise ValueError(f"SliceBatch: item {k}: the input must be contiguous (runs read it in place)")
            if (not torch.is_tensor(R) or R.device != x.device or R.dtype not in (torch.float32, torch.float64)
                    or R.ndim != 3 or tuple(R.shape[-2:]) != (3, 3) or not R.is_contiguous()):
                raise ValueError(f"SliceBatch: item {k}: rotations must be a contiguous (n, 3, 3) float32 or "
                                 f"float64 tensor on {x.device} (runs re-read it in place)")
            Rp = torch.empty(R.shape, dtype=torch.float32, device=x.device)
            self._rot.append((R, Rp))
            it = items[k]
            if kind == "extract":
                D, H, Wr = x.shape
                img = tuple(shapes[k]) if shapes is not None else (D, H, 2 * (Wr - 1))
                out = torch.empty((Rp.shape[0], img[1], _rfft_width(img[2])), dtype=x.dtype, device=x.device)
                it.vol_d, it.vol_h, it.vol_wr = D, H, Wr
                it.img_d, it.img_h, it.img_w = img
                it.output = out.data_ptr()
                self.outputs.append(out)
            else:
                D, H, W = (int(v) for v in shapes[k])
                Wr = _rfft_width(W)
                if tuple(x.shape[-2:]) != (H, Wr) or x.shape[0] != Rp.shape[0]:
                    raise ValueError(f"SliceBatch: item {k}: slices {tuple(x.shape)} do not match "
                                     f"({Rp.shape[0]}, {H}, {Wr})")
                vol = torch.empty((D, H, Wr), dtype=x.dtype, device=x.device)
                w = torch.empty((D, H, Wr), dtype=torch.float32, device=x.device)
                it.vol_d, it.vol_h, it.vol_wr = D, H, Wr
                it.img_d, it.img_h, it.img_w = D, H, W
                it.output = vol.data_ptr()
                it.weights = w.data_ptr()
                self.outputs.append((vol, w))
            it.input = x.data_ptr()
            it.rot_zyx = Rp.data_ptr()
            it.n_rot = Rp.shape[0]
            it.fftfreq_max = _fmax_arg(fftfreq_max)
            it.x_tolerance = float(x_tolerance)
            self._keep += [x, Rp]
        self.device = inputs[0].device
        self._prepare()
        handle = ctypes.c_void_p()
        kind_id = _lib.TT_BATCH_EXTRACT if kind == "extract" else _lib.TT_BATCH_INSERT
        _lib.call("tt_batch_create", kind_id, ctypes.cast(items, ctypes.c_void_p), len(inputs), int(bool(autotune)),
                  _lib.stream_of(inputs[0]), ctypes.byref(handle))
        self._handle = handle
        self.tile_shapes = []
        for k in range(len(inputs)):
            b = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int()]
            _lib.call("tt_batch_tile", self._handle, k, *[ctypes.byref(v) for v in b])
            self.tile_shapes.append(tuple(v.value for v in b))

    def _prepare(self):
        """Cast, clamp and flip every item's rotations into its bound buffer
        (slice_extraction/_extract_central_slices_rfft_3d.py:147-154,
        slice_insertion/_insert_central_slices_rfft_3d.py:16,66-67)."""
        stream = _lib.stream_of(self._keep[0])
        for R, Rp in self._rot:
            if R.shape[0] == 0:
                continue
            fn = "tt_prepare_rotations_f32" if R.dtype == torch.float32 else "tt_prepare_rotations_f64"
            _lib.call(fn, _lib.ptr(R), R.shape[0], self._flip, self._rot_tol, _lib.ptr(Rp), stream)

    def run(self):
        """Launch the batch on the current stream (no allocation: graph-capturable),
        reading the current contents of the bound inputs and rotations."""
        self._prepare()
        _lib.call("tt_batch_run", self._handle, _lib.stream_of(self._keep[0]))
        return self.outputs

    def close(self):
        if self._handle is not None:
            torch.cuda.synchronize(self.device)
            _lib.load().tt_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
