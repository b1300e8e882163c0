"""Batched extraction and insertion over volumes of uneven box size and
rotation count: one persistent work-queue launch per tile configuration,
captured in a graph, against the one-call entry points (the reference
semantics are those of a loop of one-call extract/insert calls,
slice_extraction/_extract_central_slices_rfft_3d.py:84 and
slice_insertion/_insert_central_slices_rfft_3d.py:9) and, for one item of
every box size, against the FP64 oracle."""

import math
import time

import numpy as np
import pytest
import torch

import teamtomo_amd as tt
from conftest import rel_l2
from oracle import c_oracle as C
from oracle import fourier_slice as O
from teamtomo_amd.synthetic import blob_volume, uniform_rotations, volume_rfft

pytestmark = pytest.mark.gpu


def _workload(device, n_items=32, seed=4):
    """Box sizes from {64,96,128,160,192}, rotation counts log-uniform in
    [16, 2048], uniform SO(3) rotations, seeded blob volumes."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([64, 96, 128, 160, 192], size=n_items)
    counts = np.floor(np.exp(rng.uniform(math.log(16), math.log(2048), size=n_items))).astype(int)
    vols, rots = [], []
    for k in range(n_items):
        n = int(sizes[k])
        vols.append(volume_rfft(blob_volume(n, 40, seed=100 + k, device=device)).contiguous())
        rots.append(uniform_rotations(int(counts[k]), 200 + k, device=device))
    return sizes, counts, vols, rots


def _graph(fn):
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        fn()  # warm-up outside the capture
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        fn()
    torch.cuda.synchronize()
    return g


def _time(fn, reps=5):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps * 1e3


def test_batched_extract_and_insert_match_one_call_and_replay(device):
    sizes, counts, vols, rots = _workload(device)
    print(f"batch: {len(vols)} volumes, {int(counts.sum())} rotations, sizes {sorted(set(int(s) for s in sizes))}")

    # ---- extraction
    ref = [tt.extract_central_slices_rfft_3d(v, (v.shape[0],) * 3, R) for v, R in zip(vols, rots)]
    bx = tt.SliceBatch("extract", vols, rots, autotune=True)
    out = [o.clone() for o in bx.run()]
    tiles = {int(s): bx.tile_shapes[k] for k, s in enumerate(sizes)}
    print("extract tiles per box size:", dict(sorted(tiles.items())))
    for k in range(len(vols)):
        a, b = ref[k], out[k]
        scale = float(a.abs().max())
        diff = float((a - b).abs().max())
        if bx.tile_shapes[k] == (8, 16, 16):  # the one-call engine's tile: same arithmetic
            assert torch.equal(a, b), f"item {k}: batched extraction differs from the one-call result"
        else:
            assert diff <= 1e-6 * scale, f"item {k}: {diff} vs scale {scale}"
        assert torch.equal(a == 0, b == 0)

    # ---- insertion of N(0,1) slices
    g = torch.Generator(device=device).manual_seed(5)
    slices = [torch.complex(torch.randn(o.shape, generator=g, device=device),
                            torch.randn(o.shape, generator=g, device=device)).contiguous() for o in ref]
    shapes = [(v.shape[0],) * 3 for v in vols]
    # the batch runs the brick engine's scatter: the one-call reference on the
    # same path (items with few rotations would otherwise take the binned
    # scatter, which agrees to ~1e-7 but rounds differently)
    tt.set_scatter_path("brick")
    try:
        refi = [tt.insert_central_slices_rfft_3d(s, sh, R) for s, sh, R in zip(slices, shapes, rots)]
    finally:
        tt.set_scatter_path("auto")
    bi = tt.SliceBatch("insert", slices, rots, shapes=shapes, autotune=True)
    outi = [(v.clone(), w.clone()) for v, w in bi.run()]
    tiles_i = {int(s): bi.tile_shapes[k] for k, s in enumerate(sizes)}
    print("insert tiles per box size:", dict(sorted(tiles_i.items())))
    worst = 0.0
    for k in range(len(vols)):
        (va, wa), (vb, wb) = refi[k], outi[k]
        if bi.tile_shapes[k] == (8, 8, 16):
            assert torch.equal(va, vb) and torch.equal(wa, wb), f"item {k}: batched insertion differs"
        e = max(rel_l2(vb.cpu().numpy(), va.cpu().numpy()), rel_l2(wb.cpu().numpy(), wa.cpu().numpy()))
        worst = max(worst, e)
    print(f"insert: worst rel_l2 batched vs one-call {worst:.2e}")
    assert worst < 1e-5

    # ---- the FP64 oracle, for the item with the fewest rotations of every box size
    for n in sorted(set(int(x) for x in sizes)):
        k = min((k for k in range(len(vols)) if int(sizes[k]) == n), key=lambda k: int(counts[k]))
        Rz = O.prepare_rotations(rots[k].cpu().numpy(), False, 1e-8)
        Vn = vols[k].cpu().numpy()
        e_x = rel_l2(out[k].cpu().numpy(), C.extract(Vn, (n, n, n), Rz))
        Ri = O.prepare_rotations(rots[k].cpu().numpy(), False, None)
        rV, rW = C.insert(slices[k].cpu().numpy(), (n, n, n), Ri)
        e_v = rel_l2(outi[k][0].cpu().numpy(), rV)
        e_w = rel_l2(outi[k][1].cpu().numpy(), rW)
        print(f"box {n}, item {k} ({int(counts[k])} rotations, tiles {bx.tile_shapes[k]} / {bi.tile_shapes[k]}): "
              f"vs FP64 oracle extract {e_x:.2e} insert {e_v:.2e} weights {e_w:.2e}", flush=True)
        assert e_x < 1e-5 and e_v < 1e-4 and e_w < 1e-4

    # ---- graph capture: every replay equals the eager result
    ge, gi = _graph(bx.run), _graph(bi.run)
    for rep in range(3):
        for o in bx.outputs:
            o.zero_()  # eager work between replays
        for v, w in bi.outputs:
            v.zero_()
            w.zero_()
        torch.cuda.synchronize()
        ge.replay()
        gi.replay()
        torch.cuda.synchronize()
        for k in range(len(vols)):
            assert torch.equal(bx.outputs[k], out[k]), f"replay {rep}: extraction item {k}"
            assert torch.equal(bi.outputs[k][0], outi[k][0]) and torch.equal(bi.outputs[k][1], outi[k][1]), \
                f"replay {rep}: insertion item {k}"

    # ---- timing: one graph launch vs a loop of one-call calls
    def loop():
        for v, R, s, sh in zip(vols, rots, slices, shapes):
            tt.extract_central_slices_rfft_3d(v, sh, R)
            tt.insert_central_slices_rfft_3d(s, sh, R)

    def graphs():
        ge.replay()
        gi.replay()

    t_loop, t_batch = _time(loop), _time(graphs)
    print(f"32 mixed-size volumes, extract+insert: {len(vols)} sequential one-call pairs {t_loop:.2f} ms, "
          f"batched graph replay {t_batch:.2f} ms ({t_loop / t_batch:.2f}x)")
    bx.close()
    bi.close()


def test_batch_items_without_rotations(device):
    """An item with no rotations: its extraction output is empty and its
    insertion writes a zero volume and zero weights; the other items match the
    one-call results."""
    n = 64
    vols = [volume_rfft(blob_volume(n, 20, seed=300 + k, device=device)).contiguous() for k in range(2)]
    rots = [uniform_rotations(50, 301, device=device), torch.zeros(0, 3, 3, device=device)]
    bx = tt.SliceBatch("extract", vols, rots, autotune=False)
    out = [o.clone() for o in bx.run()]
    assert out[1].shape == (0, n, n // 2 + 1)
    ref = tt.extract_central_slices_rfft_3d(vols[0], (n, n, n), rots[0])
    assert float((out[0] - ref).abs().max()) <= 1e-6 * float(ref.abs().max())
    bx.close()
    g = torch.Generator(device=device).manual_seed(6)
    slices = [torch.complex(torch.randn(ref.shape, generator=g, device=device),
                            torch.randn(ref.shape, generator=g, device=device)).contiguous(),
              torch.zeros((0, n, n // 2 + 1), dtype=torch.complex64, device=device)]
    for autotune in (False, True):
        bi = tt.SliceBatch("insert", slices, rots, shapes=[(n, n, n)] * 2, autotune=autotune)
        (v0, w0), (v1, w1) = bi.run()
        torch.cuda.synchronize()
        rv, rw = tt.insert_central_slices_rfft_3d(slices[0], (n, n, n), rots[0])
        assert rel_l2(v0.cpu().numpy(), rv.cpu().numpy()) < 1e-5 and rel_l2(w0.cpu().numpy(), rw.cpu().numpy()) < 1e-5
        assert not bool(v1.abs().amax()) and not bool(w1.abs().amax())
        bi.close()


def test_replay_reads_rotations_updated_in_place(device):
    """The batch re-reads its inputs and rotations on every run: after the
    rotations are overwritten in place, a graph replay gives the one-call
    results of the new rotations (extraction and insertion)."""
    n = 64
    vols = [volume_rfft(blob_volume(n, 20, seed=310 + k, device=device)).contiguous() for k in range(2)]
    rots = [uniform_rotations(40, 311, device=device), uniform_rotations(70, 312, device=device)]
    bx = tt.SliceBatch("extract", vols, rots, autotune=False)
    g = torch.Generator(device=device).manual_seed(313)
    slices = [torch.complex(torch.randn((r.shape[0], n, n // 2 + 1), generator=g, device=device),
                            torch.randn((r.shape[0], n, n // 2 + 1), generator=g, device=device)) for r in rots]
    bi = tt.SliceBatch("insert", slices, rots, shapes=[(n, n, n)] * 2, autotune=False)
    ge, gi = _graph(bx.run), _graph(bi.run)
    for rep in range(2):
        for k, r in enumerate(rots):
            r.copy_(uniform_rotations(r.shape[0], 400 + 10 * rep + k, device=device))  # in place
        ge.replay()
        gi.replay()
        torch.cuda.synchronize()
        for k in range(2):
            ref = tt.extract_central_slices_rfft_3d(vols[k], (n, n, n), rots[k])
            assert float((bx.outputs[k] - ref).abs().max()) <= 1e-6 * float(ref.abs().max()), f"replay {rep} item {k}"
            rv, rw = tt.insert_central_slices_rfft_3d(slices[k], (n, n, n), rots[k])
            v, w = bi.outputs[k]
            assert rel_l2(v.cpu().numpy(), rv.cpu().numpy()) < 1e-5 and rel_l2(w.cpu().numpy(), rw.cpu().numpy()) < 1e-5
    with pytest.raises(ValueError, match="contiguous"):
        tt.SliceBatch("extract", [vols[0]], [rots[0].transpose(1, 2)], autotune=False)
    bx.close()
    bi.close()
