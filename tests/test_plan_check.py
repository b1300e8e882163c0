"""Host-side bounds checks of the brick engine's work plans (tt_plan_check,
no GPU needed): at the full sizes of the headline (256^3 x 4096 gather and
adjoint), the 4096-slice insertion and the 61-tilt 512^3 insertion, and at
the odd, small, padded and masked shapes the GPU tests use, every unit's
brick, rotation segment and partial slot lies inside the buffers run_scatter
allocates, each brick's segments are covered exactly once, the split records
own disjoint slots, and the batch builder sizes an item's buffers (for the
default and the 8x8x8 tiles) at least as large."""

import ctypes

import pytest

from teamtomo_amd import _lib

GATHER, ADJOINT, INSERT, GATHER8, INSERT8 = 0, 1, 2, 3, 4


def _check(kind, vol, img, n_rot, fmax=-1.0):
    lib = _lib.load()
    st = (ctypes.c_int64 * 8)()
    rc = lib.tt_plan_check(kind, vol[0], vol[1], vol[2] // 2 + 1, img[0], img[1], img[2], n_rot, fmax, st)
    msg = lib.tt_last_error().decode(errors="replace") if rc else ""
    assert rc == 0, msg
    return dict(zip(["bricks", "units", "splits", "slots", "seg_bytes", "halo_bytes", "empty", "max_nseg"], list(st)))


def _tile(kind):
    return {GATHER: (8, 16, 16), ADJOINT: (16, 8, 16), INSERT: (8, 8, 16), GATHER8: (8, 8, 8), INSERT8: (8, 8, 8)}[kind]


@pytest.mark.parametrize("kind,vol,img,n_rot,fmax", [
    (GATHER, (256,) * 3, (256,) * 3, 4096, -1.0),
    (ADJOINT, (256,) * 3, (256,) * 3, 4096, -1.0),
    (INSERT, (256,) * 3, (256,) * 3, 4096, -1.0),
    (INSERT, (512,) * 3, (512,) * 3, 61, -1.0),
    (GATHER, (512,) * 3, (512,) * 3, 61, -1.0),
    (ADJOINT, (64,) * 3, (64,) * 3, 1024, -1.0),
    (INSERT, (64,) * 3, (64,) * 3, 1024, -1.0),
    (ADJOINT, (33,) * 3, (33,) * 3, 200, 0.3),
    (ADJOINT, (20,) * 3, (64,) * 3, 40, -1.0),
    (INSERT, (97,) * 3, (97,) * 3, 61, 0.35),
    (ADJOINT, (48,) * 3, (48,) * 3, 1100, -1.0),
    (GATHER, (256,) * 3, (256,) * 3, 1, -1.0),
    (ADJOINT, (256,) * 3, (256,) * 3, 0, -1.0),
    # the batch's alternative 8x8x8 tiles, at the box sizes of the batch test
    (GATHER8, (64,) * 3, (64,) * 3, 2048, -1.0),
    (INSERT8, (64,) * 3, (64,) * 3, 2048, -1.0),
    (GATHER8, (192,) * 3, (192,) * 3, 16, -1.0),
    (INSERT8, (160,) * 3, (160,) * 3, 700, -1.0),
    (INSERT, (192,) * 3, (192,) * 3, 2048, -1.0),
])
def test_brick_plans_stay_inside_their_buffers(kind, vol, img, n_rot, fmax):
    s = _check(kind, vol, img, n_rot, fmax)
    bz, by, bx = _tile(kind)
    nb = -(-(vol[0] - 1) // bz) * -(-(vol[1] - 1) // by) * -(-(vol[2] // 2) // bx)
    assert s["bricks"] == nb
    na = {GATHER: 0, ADJOINT: 2, INSERT: 3, GATHER8: 0, INSERT8: 3}[kind]
    nt = (bz + 1) * (by + 1) * (bx + 1)
    nsh = (by + 1) * (bx + 1) + bz * (bx + 1) + bz * by
    assert s["seg_bytes"] == s["slots"] * na * nt * 8
    assert s["halo_bytes"] == nb * na * nsh * 4
    assert s["units"] >= s["splits"]
    if kind in (GATHER, GATHER8):
        assert s["slots"] == 0 and s["empty"] == 0
    print(f"kind {kind} vol {vol} img {img} x {n_rot}: {s}")


def test_plan_check_rejects_bad_arguments():
    lib = _lib.load()
    st = (ctypes.c_int64 * 8)()
    assert lib.tt_plan_check(5, 64, 64, 33, 64, 64, 64, 10, -1.0, st) != 0
    assert lib.tt_plan_check(0, 1, 64, 33, 64, 64, 64, 10, -1.0, st) != 0
    assert lib.tt_plan_check(0, 64, 64, 33, 64, 64, 64, 10, -1.0, None) != 0
