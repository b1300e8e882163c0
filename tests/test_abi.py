"""The C-ABI library loads and exports every entry point the header declares
(CPU only: no compute calls are made)."""

import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "teamtomo_amd.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tt_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    from teamtomo_amd import _build, _lib

    if not os.path.exists(_lib.library_path()):
        _build.build()
    return ctypes.CDLL(_lib.library_path())


def test_header_declares_entry_points():
    names = _declared()
    for n in ("tt_extract_c64", "tt_extract_bwd_c64", "tt_insert_c64", "tt_insert_bwd_c64",
              "tt_sample3d_f32", "tt_splat3d_f32", "tt_last_error", "tt_abi_version"):
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


def test_abi_version_and_error_string(lib):
    from teamtomo_amd import _lib

    lib.tt_abi_version.restype = ctypes.c_int
    assert lib.tt_abi_version() == _lib.ABI_VERSION
    lib.tt_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.tt_last_error(), bytes)


def test_python_binding_covers_header():
    from teamtomo_amd import _lib

    declared = set(_declared()) - {"tt_last_error", "tt_abi_version"}
    assert declared == set(_lib._SIGNATURES), declared ^ set(_lib._SIGNATURES)


def test_invalid_arguments_rejected_without_gpu(lib):
    """Argument validation happens on the host before any HIP call."""
    from teamtomo_amd import _lib

    lib = _lib.load()
    st = lib.tt_extract_c64(None, 1, 4, 3, 4, 4, 4, None, 0, -1.0, 1e-8, None, None)
    assert st == 1 and b"volume dimensions" in lib.tt_last_error()
    st = lib.tt_sample3d_f32(None, 1, 4, 4, 4, 0, None, 0, 7, None, None)
    assert st == 1 and b"interpolation" in lib.tt_last_error()


def test_batch_create_rejects_bad_items_without_gpu(lib):
    """The batch entry point validates its items on the host before any HIP call."""
    import ctypes

    from teamtomo_amd import _lib

    lib = _lib.load()
    items = (_lib.BatchItem * 1)()
    items[0].vol_d, items[0].vol_h, items[0].vol_wr = 1, 4, 3
    items[0].img_d, items[0].img_h, items[0].img_w = 4, 4, 4
    h = ctypes.c_void_p()
    st = lib.tt_batch_create(0, ctypes.cast(items, ctypes.c_void_p), 1, 0, None, ctypes.byref(h))
    assert st == 1 and b"volume dimensions" in lib.tt_last_error() and not h.value
    st = lib.tt_batch_create(7, ctypes.cast(items, ctypes.c_void_p), 1, 0, None, ctypes.byref(h))
    assert st == 1 and b"kind" in lib.tt_last_error()
    items[0].vol_d = 4
    items[0].img_w = 8  # insertion needs the image shape to be the volume shape
    st = lib.tt_batch_create(1, ctypes.cast(items, ctypes.c_void_p), 1, 0, None, ctypes.byref(h))
    assert st == 1 and b"insertion" in lib.tt_last_error()
