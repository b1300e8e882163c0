"""Benchmark: central-slice extraction + adjoint (fwd + bwd) at 256^3 complex64.

One step = extract 4096 central slices from a (256, 256, 129) complex64 rfft
volume (the gather) and apply the adjoint to 4096 slices of N(0,1) complex64
gradients (the volume gradient, i.e. the backward of the extraction), exactly
as ``extract_central_slices_rfft_3d(...).backward(grad)`` runs it: gather
kernel, then the scatter kernel writing the gradient volume.  Inputs are resident in HBM
before timing; synthetic data (seeded blob volume, uniform SO(3) rotations).

Prints ONE JSON line.  Besides the contract fields it reports the roofline of
the dominant kernel (algorithmic bytes / HIP-event duration), the fused step's
roofline, the plain-torch comparator on the same GPU (``torch_same_gpu``, a
bounded sample of 512 rotations reported per slice; all 4096 with --torch)
and the FP64 oracle timed on the host CPU on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--torch | --no-torch] [--no-cpu]
                      [--copy-probe MIB]
                      [--workload slices3d|lines2d|insert3d|batch|project]

``--workload lines2d`` measures the 2-D analogue instead (central lines of a
(1024, 513) complex64 2-D rfft for 65536 rotations, forward + image
gradient) and ``--workload insert3d`` the insertion (4096 N(0,1) complex64
slices into (256, 256, 129), and 61 tilts of 512^2 into (512, 512, 257)) and
``--workload batch`` 32 volumes of mixed size as two graph-captured batched
launches against a loop of one-call extract + insert, and
``--workload project`` project_3d_to_2d / backproject_2d_to_3d end to end at
128^3 with padding 2 and 512 poses; each prints its own JSON line.  The
default is the headline above.
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 256
N_ROT = 4096
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--torch", action="store_true",
                    help="time the plain-torch comparator over all 4096 rotations instead of its default "
                         "bounded sample of 512")
    ap.add_argument("--no-torch", action="store_true",
                    help="skip the plain-torch comparator (elementwise coordinates, advanced-indexing "
                         "gather, index_add_ adjoint) on the same GPU")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--deterministic", action="store_true")
    ap.add_argument("--gather", choices=["auto", "pixel", "brick"], default="auto",
                    help="extraction gather kernel (default: auto)")
    ap.add_argument("--workload", choices=["slices3d", "lines2d", "insert3d", "batch", "project"],
                    default="slices3d",
                    help="slices3d: the headline (default); lines2d: the 2-D central-line analogue; "
                         "insert3d: central-slice insertion (4096 slices into 256^3, 61 tilts into 512^3); "
                         "batch: 32 volumes of mixed size, batched extract + insert against one-call loops; "
                         "project: project_3d_to_2d / backproject_2d_to_3d end to end (128^3, pad 2, 512 poses)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB and exit (calibrates the HBM "
                         "FETCH_SIZE / WRITE_SIZE counters against a copy of known size)")
    ap.add_argument("--write-traffic", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "FETCH_DIR_INSERT",
                                                         "WRITE_DIR_INSERT"),
                    help="only write profiles/traffic.json from four rocprofv3 --pmc output directories "
                         "(FETCH_SIZE and WRITE_SIZE runs of the default workload, then of --workload "
                         "insert3d) and print their per-dispatch tables; no GPU is used")
    return ap.parse_args()


# Launch sequence of one call of each measured entry point, in dispatch order
# (kernel names as rocprofv3 prints them, without the argument list).
CALL_KERNELS = {
    "extract": ["plane_rec_kernel", "gather_kernel<8, 16, 16, false, 1>"],
    "adjoint": ["fill_zero_kernel<unsigned int>", "band_max_kernel", "plane_rec_kernel",
                "brick_kernel<16, 8, 16, 1, false, 1>", "split_combine_kernel<16, 8, 16, 1>",
                "halo_combine_kernel<16, 8, 16, 1>", "maxabs_chunks_kernel",
                "brick_kernel<16, 8, 16, 1, false, 1>", "split_combine_kernel<16, 8, 16, 1>",
                "halo_combine_kernel<16, 8, 16, 1>"],
    "insert_256": ["fill_zero_kernel<unsigned int>", "band_max_kernel", "plane_rec_kernel",
                   "brick_kernel<8, 8, 16, 2, false, 1>", "split_combine_kernel<8, 8, 16, 2>",
                   "halo_combine_kernel<8, 8, 16, 2>", "maxabs_chunks_kernel",
                   "brick_kernel<8, 8, 16, 2, false, 1>", "split_combine_kernel<8, 8, 16, 2>",
                   "halo_combine_kernel<8, 8, 16, 2>", "mirror_kernel"],
    "insert_512_tilts": ["bin_zero_kernel", "bin_count_kernel<true>", "bin_scan_kernel", "bin_fill_kernel<true>",
                         "bin_brick_kernel<true>", "mirror_kernel"],
}
# Bytes per counter unit, from a 1 GiB device-to-device copy (--copy-probe
# 1024): FETCH_SIZE reads 524300 KiB and WRITE_SIZE 1048576 KiB per copy.
READ_BYTES_PER_FETCH_UNIT = 2.0 * 1024
WRITE_BYTES_PER_WRITE_UNIT = 1.0 * 1024


def _kernel_name(full):
    """rocprofv3's kernel name without 'void', the tt namespaces and the argument list."""
    n = full[5:] if full.startswith("void ") else full
    n = n.replace("tt::(anonymous namespace)::", "").replace("tt::", "")
    depth = 0
    for i, c in enumerate(n):
        if c == "<":
            depth += 1
        elif c == ">":
            depth -= 1
        elif c == "(" and depth == 0:
            return n[:i]
    return n


def _pmc_dispatches(d, counter):
    """[(dispatch id, kernel name, value)] of one rocprofv3 --pmc run, in dispatch order."""
    import csv

    rows = []
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                rows.append((int(r["Dispatch_Id"]), _kernel_name(r["Kernel_Name"]), float(r["Counter_Value"])))
    rows.sort()
    return rows


def _last_call(rows, seq):
    """The dispatches of the last occurrence of the launch sequence seq."""
    names = [r[1] for r in rows]
    for i in range(len(rows) - len(seq), -1, -1):
        if names[i:i + len(seq)] == seq:
            return rows[i:i + len(seq)]
    raise SystemExit(f"launch sequence {seq} not found")


def write_traffic(dirs):
    """profiles/traffic.json from FETCH_SIZE / WRITE_SIZE runs (see --write-traffic), and
    the per-dispatch tables of the last call of each entry point on stdout."""
    fh, wh, fi, wi = dirs
    runs = {"extract": (fh, wh), "adjoint": (fh, wh), "insert_256": (fi, wi), "insert_512_tilts": (fi, wi)}
    out = {"_note": "HBM-side bytes per launch sequence (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, KiB->bytes) "
                    "from rocprofv3 --pmc runs of bench.py (--write-traffic) on the kernel sources of "
                    "source_digest; see extract_adjoint_256c64_4096rot.hbm_traffic.md and "
                    "insert_256c64_4096rot_512c64_61tilts.hbm_traffic.md",
           "source_digest": kernel_source_digest()}
    per = {}
    for key, (fd, wd) in runs.items():
        seq = CALL_KERNELS[key]
        fr = _last_call(_pmc_dispatches(fd, "FETCH_SIZE"), seq)
        wr = _last_call(_pmc_dispatches(wd, "WRITE_SIZE"), seq)
        total = 0.0
        print(f"\n{key}\n\n| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | read MB | written MB |\n|---|---|---|---|---|")
        for (_, name, f), (_, _, w) in zip(fr, wr):
            rb, wb = f * READ_BYTES_PER_FETCH_UNIT, w * WRITE_BYTES_PER_WRITE_UNIT
            total += rb + wb
            per[f"{key}/{name}"] = per.get(f"{key}/{name}", 0.0) + rb + wb
            print(f"| {name} | {f:.0f} | {w:.0f} | {rb / 1e6:.1f} | {wb / 1e6:.1f} |")
        out[key] = total
        print(f"\n{key}: {total / 1e6:.0f} MB per call")
    out["per_kernel"] = per
    with open(os.path.join(ROOT, "profiles", "traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"\nwrote profiles/traffic.json (source digest {out['source_digest']})")


def make_inputs(device):
    from teamtomo_amd.synthetic import blob_volume, uniform_rotations, volume_rfft

    vol = blob_volume(N, 200, seed=2, device=device)
    V = volume_rfft(vol).contiguous()
    R = uniform_rotations(N_ROT, seed=0, device=device)
    g = torch.Generator(device=device).manual_seed(3)
    gS = torch.complex(
        torch.randn((N_ROT, N, N // 2 + 1), generator=g, device=device),
        torch.randn((N_ROT, N, N // 2 + 1), generator=g, device=device),
    ).contiguous()
    return V, R, gS


class Step:
    """extract + adjoint through the C ABI, with HIP events around each kernel.

    The adjoint writes the gradient volume (TT_FLAG_OVERWRITE), as the
    autograd backward does, so no separate zero fill is needed."""

    def __init__(self, V, R, gS, deterministic=False):
        from teamtomo_amd import _lib
        from teamtomo_amd.fourier_slice import _prepare_rotations

        self.lib = _lib
        self.V, self.gS = V, gS
        self.Rz = _prepare_rotations(R, V.device, False, 1e-8).reshape(-1, 3, 3).contiguous()
        self.n = self.Rz.shape[0]
        self.out = torch.empty((self.n, N, N // 2 + 1), dtype=torch.complex64, device=V.device)
        self.gV = torch.empty_like(V)
        self.flags = _lib.TT_FLAG_OVERWRITE | (_lib.TT_FLAG_DETERMINISTIC if deterministic else 0)
        self.stream = _lib.stream_of(V)
        self.events = []

    def __call__(self, timed=False):
        L, V = self.lib, self.V
        s = torch.cuda.current_stream()
        if timed:
            e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e[0].record(s)
        L.call("tt_extract_c64", L.ptr(V), N, N, N // 2 + 1, N, N, N, L.ptr(self.Rz), self.n, -1.0, 1e-8,
               L.ptr(self.out), self.stream)
        if timed:
            e[1].record(s)
        L.call("tt_extract_bwd_c64", L.ptr(self.gS), L.ptr(V), N, N, N // 2 + 1, N, N, N, L.ptr(self.Rz),
               self.n, -1.0, 1e-8, L.ptr(self.gV), None, self.flags, self.stream)
        if timed:
            e[2].record(s)
            self.events.append(e)


def algorithmic_bytes():
    P = N * (N // 2 + 1)
    V = N * N * (N // 2 + 1)
    extract = N_ROT * (8 * P + 36) + 8 * V
    adjoint = N_ROT * (8 * P + 36) + 8 * V
    return extract, adjoint


def _torch_coords(image_shape, R, device):
    """Sampling coordinates of the reference op sequence
    (slice_extraction/_extract_central_slices_rfft_3d.py:147-210, _dft_utils.py:5-72),
    in float32 with elementwise ops (the rotation is written out as products and sums)."""
    d, h, w = image_shape
    R = R.to(torch.float32)
    R = torch.where(torch.abs(R) < 1e-8, torch.zeros((), device=device), R)
    R = torch.flip(R, dims=(-2, -1))
    fy = torch.fft.fftshift(torch.fft.fftfreq(h, device=device))[:, None]
    fx = torch.fft.rfftfreq(w, device=device)[None, :]
    q = [R[:, k, 1, None, None] * fy + R[:, k, 2, None, None] * fx for k in range(3)]
    r11, r12 = R[:, 2, 2, None, None], R[:, 2, 1, None, None]
    tie = torch.where(torch.abs(r11) <= 1e-7, r12 < 0, r11 < 0)
    conj = (q[2] < -1e-8) | ((torch.abs(q[2]) < 1e-8) & tie)
    sgn = 1.0 - 2.0 * conj.to(torch.float32)
    c = torch.stack([(sgn * q[0] + 0.5) / (1 - 1 / d) * (d - 1), (sgn * q[1] + 0.5) / (1 - 1 / h) * (h - 1),
                     sgn * q[2] * 2 * (w // 2)], dim=-1)
    return c, conj


def _torch_corners(c, shape):
    dims = torch.tensor(shape, device=c.device, dtype=torch.float32)
    inside = torch.all((c >= 0) & (c <= dims - 1), dim=-1)
    cc = torch.where(inside[..., None], c, torch.zeros((), device=c.device))
    lo = torch.floor(cc)
    t = cc - lo
    lo = lo.long()
    for q in range(8):
        dz, dy, dx = q >> 2, (q >> 1) & 1, q & 1
        w = (t[..., 0] if dz else 1 - t[..., 0]) * (t[..., 1] if dy else 1 - t[..., 1]) * \
            (t[..., 2] if dx else 1 - t[..., 2])
        iz = torch.clamp(lo[..., 0] + dz, max=shape[0] - 1)
        iy = torch.clamp(lo[..., 1] + dy, max=shape[1] - 1)
        ix = torch.clamp(lo[..., 2] + dx, max=shape[2] - 1)
        yield (iz * shape[1] + iy) * shape[2] + ix, w * inside


def torch_extract(V, image_shape, R):
    """Plain-torch trilinear gather: 8 advanced-indexing corner reads per sample."""
    c, conj = _torch_coords(image_shape, R, V.device)
    Vf = V.reshape(-1)
    s = torch.zeros(c.shape[:-1], dtype=V.dtype, device=V.device)
    for idx, w in _torch_corners(c, V.shape):
        s += w * Vf[idx]
    return torch.where(conj, torch.conj(s), s)


def torch_adjoint(gS, gV, image_shape, R):
    """Plain-torch adjoint: index_add_ of w * conj-folded gradients at the 8 corners."""
    c, conj = _torch_coords(image_shape, R, gV.device)
    g = torch.where(conj, torch.conj(gS), gS)
    gr = torch.view_as_real(g.resolve_conj()).reshape(-1, 2)
    out = torch.view_as_real(gV).view(-1, 2)
    for idx, w in _torch_corners(c, gV.shape):
        out.index_add_(0, idx.reshape(-1), gr * w.reshape(-1, 1))


def torch_insert(S, vol, wts, R):
    """Plain-torch insertion (slice_insertion/_insert_central_slices_rfft_3d.py:
    9-106, :289-406): conj where x' < 0, x = 0 / y < 0 pixels skipped, floor/ceil
    corners, index_add_ of w * value and w, and the Friedel mirror of every
    x = 0 corner but DC."""
    d, h, wr = vol.shape
    w = 2 * (wr - 1)
    R = torch.flip(R.to(torch.float32), dims=(-2, -1))
    fy = torch.fft.fftshift(torch.fft.fftfreq(h, device=S.device))[:, None]
    fx = torch.fft.rfftfreq(w, device=S.device)[None, :]
    q = [R[:, k, 1, None, None] * fy + R[:, k, 2, None, None] * fx for k in range(3)]
    conj = q[2] < 0
    sgn = 1.0 - 2.0 * conj.to(torch.float32)
    c = torch.stack([(sgn * q[0] + 0.5) / (1 - 1 / d) * (d - 1), (sgn * q[1] + 0.5) / (1 - 1 / h) * (h - 1),
                     sgn * q[2] * 2 * (wr - 1)], dim=-1)
    keep = ~((fx == 0) & (fy < 0)).expand(S.shape[-2:])
    val = torch.where(conj, torch.conj(S), S)
    dims = torch.tensor([d, h, wr], device=S.device, dtype=torch.float32)
    ok = keep[None] & torch.all((c >= 0) & (c <= dims - 1), dim=-1)
    c, val = c[ok], val[ok]
    lo, hi = torch.floor(c), torch.ceil(c)
    t = c - lo
    lo, hi = lo.long(), hi.long()
    vr = torch.view_as_real(vol).view(-1, 2)
    wf = wts.view(-1)
    vv = torch.view_as_real(val.resolve_conj())
    for k in range(8):
        dz, dy, dx = k >> 2, (k >> 1) & 1, k & 1
        iz, iy, ix = (hi if dz else lo)[:, 0], (hi if dy else lo)[:, 1], (hi if dx else lo)[:, 2]
        wk = (t[:, 0] if dz else 1 - t[:, 0]) * (t[:, 1] if dy else 1 - t[:, 1]) * (t[:, 2] if dx else 1 - t[:, 2])
        idx = (iz * h + iy) * wr + ix
        vr.index_add_(0, idx, vv * wk[:, None])
        wf.index_add_(0, idx, wk)
        m = (ix == 0) & ~((iz == d // 2) & (iy == h // 2))  # the x = 0 Friedel mirror
        midx = (((d - iz[m]) % d) * h + (h - iy[m]) % h) * wr
        vr.index_add_(0, midx, torch.stack([vv[m, 0], -vv[m, 1]], -1) * wk[m, None])
        wf.index_add_(0, midx, wk[m])


def time_torch_comparator(V, R, gS, steps=2, chunk=256, n_rot=512):
    """Plain-torch extract + adjoint on the same GPU over the first n_rot
    rotations, chunked over rotations; returns slices/s."""

    def step():
        out = []
        gV = torch.zeros_like(V)
        for r0 in range(0, n_rot, chunk):
            r1 = min(r0 + chunk, n_rot)
            out.append(torch_extract(V, (N, N, N), R[r0:r1]))
            torch_adjoint(gS[r0:r1], gV, (N, N, N), R[r0:r1])
        return out, gV

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return n_rot / dt, dt * 1e3, chunk


def cpu_baseline(V, R, gS, step, min_s=10.0, chunk=256):
    """FP64 C oracle (OpenMP over rotations) on the host CPU: extract + adjoint of
    the rotations in chunks of `chunk` until `min_s` seconds have passed or all
    N_ROT are done (a bounded sample of the workload), reported per slice.
    The oracle's results for the first chunk also check the GPU's on the same
    inputs at full size: the slices of the last timed step, and an adjoint of
    that chunk's gradients."""
    from oracle import c_oracle as C
    from oracle import fourier_slice as O

    Vn = V.cpu().numpy().astype(np.complex128)
    Rn = O.prepare_rotations(R.cpu().numpy(), False, 1e-8)
    g = gS.cpu().numpy()
    C.extract(Vn, (N, N, N), Rn[:2])  # load + first-touch outside the timing
    done, dt, first = 0, 0.0, None
    while done < N_ROT and dt < min_s:
        sl = slice(done, min(done + chunk, N_ROT))
        t0 = time.perf_counter()
        S = C.extract(Vn, (N, N, N), Rn[sl])
        G = C.extract_adjoint(g[sl], Vn.shape, (N, N, N), Rn[sl])
        dt += time.perf_counter() - t0
        if first is None:
            first = (sl, S, G)
        done = sl.stop
        del S, G
    return {"value": done / dt, "unit": "slices/s", "cores": C.threads(), "kind": "port",
            "sample": f"FP64 C oracle (oracle/c/tt_oracle.c, OpenMP), extract+adjoint of {done} of the "
                      f"{N_ROT} rotations at {N}^3 in {dt:.1f} s",
            "parity": parity_check(step, *first)}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def torch_cpu_baseline(min_s=5.0):
    """The plain-torch restatement (torch_extract / torch_adjoint above) on the
    host CPU with torch's intra-op threads: a 64^3 N(0,1) volume with 100
    rotations in full and the 256^3 workload on rotation chunks of 16
    until `min_s` seconds have passed, both reported per slice; plus the FP64
    C oracle on the 64^3 case."""
    from oracle import c_oracle as C
    from oracle import fourier_slice as O
    from teamtomo_amd.synthetic import blob_volume, uniform_rotations, volume_rfft

    cpu = torch.device("cpu")
    out = {"threads": torch.get_num_threads(), "cpu": _cpu_model()}
    n1 = 64
    V1 = volume_rfft(torch.randn((n1,) * 3, generator=torch.Generator().manual_seed(1))).contiguous()
    R1 = uniform_rotations(100, seed=1, device=cpu)
    g1 = torch.complex(torch.randn((100, n1, n1 // 2 + 1), generator=torch.Generator().manual_seed(2)),
                       torch.randn((100, n1, n1 // 2 + 1), generator=torch.Generator().manual_seed(3)))
    torch_extract(V1, (n1,) * 3, R1[:2])
    t0 = time.perf_counter()
    torch_extract(V1, (n1,) * 3, R1)
    torch_adjoint(g1, torch.zeros_like(V1), (n1,) * 3, R1)
    dt = time.perf_counter() - t0
    out["vol64_100rot"] = {"value": 100 / dt, "unit": "slices/s", "seconds": dt}
    Rn = O.prepare_rotations(R1.numpy(), False, 1e-8)
    Vn = V1.numpy().astype(np.complex128)
    t0 = time.perf_counter()
    C.extract(Vn, (n1,) * 3, Rn)
    C.extract_adjoint(g1.numpy(), Vn.shape, (n1,) * 3, Rn)
    dt = time.perf_counter() - t0
    out["vol64_100rot_c_oracle"] = {"value": 100 / dt, "unit": "slices/s", "seconds": dt, "threads": C.threads()}
    V2 = volume_rfft(blob_volume(N, 200, seed=2, device=cpu)).contiguous()
    R2 = uniform_rotations(N_ROT, seed=0, device=cpu)
    gen = torch.Generator().manual_seed(4)
    done, dt = 0, 0.0
    while dt < min_s and done < N_ROT:
        k = 16
        gS = torch.complex(torch.randn((k, N, N // 2 + 1), generator=gen), torch.randn((k, N, N // 2 + 1), generator=gen))
        t0 = time.perf_counter()
        torch_extract(V2, (N, N, N), R2[done:done + k])
        torch_adjoint(gS, torch.zeros_like(V2), (N, N, N), R2[done:done + k])
        dt += time.perf_counter() - t0
        done += k
    out["vol256_4096rot_sample"] = {"value": done / dt, "unit": "slices/s", "rotations": done, "seconds": dt}
    return out


def parity_check(step, sl, S_ref, G_ref):
    """Relative L2 and max |error| (over max |reference|) of the GPU's slices
    and volume gradient against the FP64 oracle for the rotations in `sl`."""
    L = step.lib
    n = sl.stop - sl.start
    gS = step.gS[sl].contiguous()
    Rz = step.Rz[sl].contiguous()
    gV = torch.empty_like(step.V)
    L.call("tt_extract_bwd_c64", L.ptr(gS), L.ptr(step.V), N, N, N // 2 + 1, N, N, N, L.ptr(Rz), n, -1.0, 1e-8,
           L.ptr(gV), None, step.flags, step.stream)
    torch.cuda.synchronize()

    def cmp(a, b):
        den = float(np.linalg.norm(b))
        return {"rel_l2": float(np.linalg.norm(a - b)) / den, "max_abs_over_max": float(np.max(np.abs(a - b))) /
                float(np.max(np.abs(b)))}

    return {"rotations": [sl.start, sl.stop], "slices": cmp(step.out[sl].cpu().numpy(), S_ref),
            "volume_gradient": cmp(gV.cpu().numpy(), G_ref),
            "bar": "rel_l2 <= 1e-5 (slices), <= 1e-4 (gradient)"}


L2D_N = 1024      # 2-D image side: rfft (1024, 513)
L2D_ROT = 65536   # rotations (lines) per step


def run_lines2d(args, device):
    """Central lines of a 2-D rfft, forward + image gradient, as
    ``extract_central_slices_rfft_2d(...).backward(grad)`` runs them: the line
    gather, then a zero fill of the image gradient and the adjoint splat."""
    from teamtomo_amd import _lib

    n, wr = L2D_N, L2D_N // 2 + 1
    g = torch.Generator(device=device).manual_seed(21)
    img = torch.randn((n, n), generator=g, device=device)
    F = torch.fft.fftshift(torch.fft.rfftn(torch.fft.fftshift(img, dim=(-2, -1)), dim=(-2, -1)), dim=(-2,))
    F = F.to(torch.complex64).contiguous()
    th = torch.rand(L2D_ROT, generator=g, device=device, dtype=torch.float64) * 2 * np.pi
    R = torch.stack([torch.cos(th), -torch.sin(th), torch.sin(th), torch.cos(th)], -1).reshape(-1, 2, 2)
    Ryx = torch.flip(R.to(torch.float32), dims=(-2, -1)).contiguous()
    gL = torch.complex(torch.randn((L2D_ROT, wr), generator=g, device=device),
                       torch.randn((L2D_ROT, wr), generator=g, device=device)).contiguous()
    out = torch.empty((L2D_ROT, wr), dtype=torch.complex64, device=device)
    gI = torch.empty_like(F)
    st = _lib.stream_of(F)
    events = []

    def step(timed=False):
        s = torch.cuda.current_stream()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)] if timed else None
        if timed:
            e[0].record(s)
        _lib.call("tt_extract2d_c64", _lib.ptr(F), n, wr, n, n, _lib.ptr(Ryx), L2D_ROT, -1.0, _lib.ptr(out), st)
        if timed:
            e[1].record(s)
        gI.zero_()
        _lib.call("tt_extract2d_bwd_c64", _lib.ptr(gL), _lib.ptr(F), n, wr, n, n, _lib.ptr(Ryx), L2D_ROT, -1.0,
                  _lib.ptr(gI), None, st)
        if timed:
            e[2].record(s)
            events.append(e)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        step(timed=True)
    stop.record()
    torch.cuda.synchronize()
    ms_step = start.elapsed_time(stop) / args.steps
    fwd_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in events]))
    adj_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in events]))
    # algorithmic bytes: 8 B per line pixel + 16 B per rotation, plus the image read
    # (forward) or written (adjoint) once
    fwd_b = L2D_ROT * (8 * wr + 16) + 8 * n * wr
    adj_b = L2D_ROT * (8 * wr + 16) + 8 * n * wr
    kernels = {"forward": {"ms": fwd_ms, "bytes": fwd_b, "GBs": fwd_b / fwd_ms / 1e6,
                           "launch": "tt_extract2d_c64: line_fwd_kernel"},
               "adjoint": {"ms": adj_ms, "bytes": adj_b, "GBs": adj_b / adj_ms / 1e6,
                           "launch": "image-gradient zero fill + tt_extract2d_bwd_c64: line_bwd_kernel"}}
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    line = {
        "metric": "lines/s fwd+adjoint of 2-D central lines at 1024^2 c64; % HBM roofline",
        "value": L2D_ROT / (ms_step / 1e3), "unit": "lines/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64",
        "data": "synthetic (seeded N(0,1) 1024^2 image, uniform in-plane angles, N(0,1) c64 gradients)",
        "config": {"workload": "extract_central_slices_rfft_2d fwd+adjoint, rfft (1024, 513) c64, 65536 rotations",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": kernels[dom]["GBs"] / HBM_PEAK_GBS, "traffic": None},
        "kernels": kernels,
    }
    if not args.no_cpu:
        from oracle import slice_2d as O2

        Fn = F.cpu().numpy().astype(np.complex128)
        Rn = Ryx[:512].cpu().numpy()
        gn = gL[:512].cpu().numpy()
        O2.extract_central_lines(Fn, (n, n), Rn[:2])
        t0 = time.perf_counter()
        ref_l = O2.extract_central_lines(Fn, (n, n), Rn)
        ref_g = O2.extract_central_lines_adjoint(gn, Fn.shape, (n, n), Rn)
        dt = time.perf_counter() - t0
        # the same 512 lines on the GPU: the timed step's forward output, and
        # an image gradient of those lines alone
        gI1 = torch.zeros_like(F)
        _lib.call("tt_extract2d_bwd_c64", _lib.ptr(gL[:512].contiguous()), _lib.ptr(F), n, wr, n, n,
                  _lib.ptr(Ryx[:512].contiguous()), 512, -1.0, _lib.ptr(gI1), None, st)
        torch.cuda.synchronize()

        def cmp(a, b):
            return {"rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                    "max_abs_over_max": float(np.max(np.abs(a - b)) / np.max(np.abs(b)))}

        line["cpu_baseline"] = {"value": 512 / dt, "unit": "lines/s", "cores": 1, "kind": "port",
                                "sample": f"FP64 NumPy oracle (oracle/slice_2d.py), forward + image adjoint of "
                                          f"512 of the {L2D_ROT} lines in {dt:.2f} s",
                                "parity": {"rotations": [0, 512], "lines": cmp(out[:512].cpu().numpy(), ref_l),
                                           "image_gradient": cmp(gI1.cpu().numpy(), ref_g),
                                           "bar": "rel_l2 <= 1e-5"}}
    print(json.dumps(line), flush=True)


def copy_bandwidth(device, mib=1024, reps=5):
    """Measured HBM copy rate (GB/s, read + write bytes) of a 1 GiB
    device-to-device copy: the practical ceiling next to the 8 TB/s peak."""
    src = torch.ones(mib << 18, device=device)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2 * (mib << 20) / ms / 1e6


def _code_only(text):
    """C++ source without its comments and with whitespace runs collapsed, so
    that edits to comments alone leave the digest unchanged."""
    import re

    text = re.sub(r"//[^\n]*|/\*.*?\*/", " ", text, flags=re.S)
    return " ".join(text.split())


def kernel_source_digest():
    """sha256 (first 16 hex digits) over the code of the HIP sources the
    library is built from (comments and layout excluded).
    profiles/traffic.json records the digest of the sources its counters were
    measured on; its figures are reported only while that code is unchanged."""
    h = hashlib.sha256()
    d = os.path.join(ROOT, "teamtomo_amd", "csrc")
    for name in sorted(os.listdir(d)):
        if name.endswith((".hip", ".h")):
            h.update(name.encode())
            with open(os.path.join(d, name), encoding="utf-8") as f:
                h.update(_code_only(f.read()).encode())
    return h.hexdigest()[:16]


def measured_traffic(key):
    """HBM-side bytes per launch sequence of `key` from profiles/traffic.json,
    or (None, reason) when absent or measured on other kernel sources."""
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(tpath):
        return None, "no profiles/traffic.json"
    with open(tpath) as f:
        t = json.load(f)
    if t.get("source_digest") != kernel_source_digest():
        return None, "profiles/traffic.json was measured on other kernel sources"
    v = t.get(key)
    return v, (None if v is not None else f"no {key} entry in profiles/traffic.json")


def run_insert3d(args, device):
    """Central-slice insertion as ``insert_central_slices_rfft_3d`` runs it
    (volume and weights written): 4096 N(0,1) complex64 slices with uniform
    rotations into (256, 256, 129), and 61 tilts (-60..60 step 2 about y) of
    512^2 into (512, 512, 257)."""
    from teamtomo_amd import _lib
    from teamtomo_amd.fourier_slice import _prepare_rotations, _scatter_flags
    from teamtomo_amd.synthetic import uniform_rotations

    def case(n, R, seed):
        g = torch.Generator(device=device).manual_seed(seed)
        Rz = _prepare_rotations(R, device, False, None).reshape(-1, 3, 3).contiguous()
        k = Rz.shape[0]
        S = torch.complex(torch.randn((k, n, n // 2 + 1), generator=g, device=device),
                          torch.randn((k, n, n // 2 + 1), generator=g, device=device)).contiguous()
        vol = torch.empty((n, n, n // 2 + 1), dtype=torch.complex64, device=device)
        w = torch.empty((n, n, n // 2 + 1), dtype=torch.float32, device=device)
        st = _lib.stream_of(S)

        def run(S=S, Rz=Rz, k=k):
            _lib.call("tt_insert_c64", _lib.ptr(S), n, n, n, _lib.ptr(Rz), k, -1.0, _lib.ptr(vol), _lib.ptr(w),
                      _scatter_flags(), st)

        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        P, V = n * (n // 2 + 1), n * n * (n // 2 + 1)
        nbytes = k * (8 * P + 36) + 12 * V  # slices + rotations read, volume + weights written
        return {"ms": ms, "bytes": nbytes, "GBs": nbytes / ms / 1e6, "slices": k, "volume": [n, n, n // 2 + 1],
                "slices_per_s": k / (ms / 1e3)}, (S, Rz, vol, w, run)

    c2, (S2, Rz2, vol2, w2, run2) = case(N, uniform_rotations(N_ROT, seed=0, device=device), 3)
    t = np.deg2rad(np.arange(-60, 61, 2, dtype=np.float64))
    Ry = np.zeros((t.size, 3, 3), dtype=np.float32)
    Ry[:, 0, 0], Ry[:, 0, 2], Ry[:, 1, 1], Ry[:, 2, 0], Ry[:, 2, 2] = np.cos(t), np.sin(t), 1.0, -np.sin(t), np.cos(t)
    c5, _ = case(512, torch.from_numpy(Ry).to(device), 5)
    torch_line = None
    if args.torch:
        def torch_step():
            vt = torch.zeros_like(vol2)
            wt = torch.zeros_like(w2)
            for r0 in range(0, N_ROT, 256):
                torch_insert(S2[r0:r0 + 256], vt, wt, Rz2[r0:r0 + 256].flip(dims=(-2, -1)))
            return vt, wt

        vt, wt = torch_step()
        run2()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        torch_step()
        torch.cuda.synchronize()
        tms = (time.perf_counter() - t0) * 1e3
        rel_v = float(torch.linalg.vector_norm(vt - vol2) / torch.linalg.vector_norm(vt))
        rel_w = float(torch.linalg.vector_norm(wt - w2) / torch.linalg.vector_norm(wt))
        torch_line = {"value": N_ROT / (tms / 1e3), "unit": "slices/s", "ms_per_step": tms, "rotation_chunk": 256,
                      "speedup": tms / c2["ms"], "rel_l2_vs_ours": {"volume": rel_v, "weights": rel_w}}
    line = {
        "metric": "slices/s central-slice insertion at 256^3 c64 (4096 slices); % HBM roofline",
        "value": c2["slices_per_s"], "unit": "slices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": c2["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64",
        "data": "synthetic (N(0,1) c64 slices, uniform SO(3) rotations; 61 tilts about y for the 512^3 case)",
        "config": {"workload": "insert_central_slices_rfft_3d, 4096 slices (256, 129) into (256, 256, 129) c64 + "
                               "f32 weights", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "insert", "achieved": c2["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": c2["GBs"] / HBM_PEAK_GBS, "traffic": None},
        "kernels": {"insert_256": c2, "insert_512_tilts": dict(c5, frac=c5["GBs"] / HBM_PEAK_GBS)},
    }
    for key, c in (("insert_256", c2), ("insert_512_tilts", c5)):
        tr, why = measured_traffic(key)
        line["kernels"][key]["traffic"] = tr
        line["kernels"][key]["traffic_per_algorithmic_byte"] = tr / c["bytes"] if tr else None
        if why:
            line["kernels"][key]["traffic_note"] = why
    line["roofline"]["traffic"] = line["kernels"]["insert_256"]["traffic"]
    if torch_line is not None:
        line["torch_same_gpu"] = torch_line
    if not args.no_cpu:
        from oracle import c_oracle as C
        from oracle import fourier_slice as O

        Sn = S2.cpu().numpy()
        Rn = O.prepare_rotations(Rz2.cpu().numpy(), True, None)
        C.insert(Sn[:2], (N, N, N), Rn[:2])
        done, dt, first = 0, 0.0, None
        while done < N_ROT and dt < 10.0:
            sl = slice(done, min(done + 256, N_ROT))
            t0 = time.perf_counter()
            res = C.insert(Sn[sl], (N, N, N), Rn[sl])
            dt += time.perf_counter() - t0
            if first is None:
                first = (sl, res)
            done = sl.stop
        sl, (rv, rw) = first
        gv = torch.empty_like(vol2)
        gw = torch.empty_like(w2)
        k = sl.stop - sl.start
        _lib.call("tt_insert_c64", _lib.ptr(S2[sl].contiguous()), N, N, N, _lib.ptr(Rz2[sl].contiguous()), k, -1.0,
                  _lib.ptr(gv), _lib.ptr(gw), _scatter_flags(), _lib.stream_of(S2))
        torch.cuda.synchronize()

        def cmp(a, b):
            return {"rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                    "max_abs_over_max": float(np.max(np.abs(a - b)) / np.max(np.abs(b)))}

        line["cpu_baseline"] = {
            "value": done / dt, "unit": "slices/s", "cores": C.threads(), "kind": "port",
            "sample": f"FP64 C oracle (oracle/c/tt_oracle.c, OpenMP), insertion of {done} of the {N_ROT} slices "
                      f"at {N}^3 in {dt:.1f} s",
            "parity": {"rotations": [sl.start, sl.stop], "volume": cmp(gv.cpu().numpy(), rv),
                       "weights": cmp(gw.cpu().numpy(), rw), "bar": "rel_l2 <= 1e-4"}}
    print(json.dumps(line), flush=True)


def run_batch(args, device):
    """32 volumes with box sizes from {64, 96, 128, 160, 192} and rotation
    counts log-uniform in [16, 2048] (seed 4): extraction and insertion of
    every volume as two graph-captured SliceBatch launches, against a loop of
    one-call extract + insert per volume."""
    import math

    import teamtomo_amd as tt
    from teamtomo_amd.synthetic import blob_volume, uniform_rotations, volume_rfft

    rng = np.random.default_rng(4)
    sizes = rng.choice([64, 96, 128, 160, 192], size=32)
    counts = np.floor(np.exp(rng.uniform(math.log(16), math.log(2048), size=32))).astype(int)
    vols = [volume_rfft(blob_volume(int(n), 40, seed=100 + k, device=device)).contiguous() for k, n in enumerate(sizes)]
    rots = [uniform_rotations(int(c), 200 + k, device=device) for k, c in enumerate(counts)]
    shapes = [(int(n),) * 3 for n in sizes]
    g = torch.Generator(device=device).manual_seed(5)
    slices = [torch.complex(torch.randn((int(c), int(n), int(n) // 2 + 1), generator=g, device=device),
                            torch.randn((int(c), int(n), int(n) // 2 + 1), generator=g, device=device)).contiguous()
              for n, c in zip(sizes, counts)]
    bx = tt.SliceBatch("extract", vols, rots, autotune=True)
    bi = tt.SliceBatch("insert", slices, rots, shapes=shapes, autotune=True)

    def capture(fn):
        graph = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            fn()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        with torch.cuda.graph(graph):
            fn()
        torch.cuda.synchronize()
        return graph

    ge, gi = capture(bx.run), capture(bi.run)

    def batched():
        ge.replay()
        gi.replay()

    def loop():
        for v, R, sl, sh in zip(vols, rots, slices, shapes):
            tt.extract_central_slices_rfft_3d(v, sh, R)
            tt.insert_central_slices_rfft_3d(sl, sh, R)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    ms_b, ms_l = timed(batched), timed(loop)
    # parity of the batched outputs with the one-call results
    worst = 0.0
    for k in range(len(vols)):
        a = tt.extract_central_slices_rfft_3d(vols[k], shapes[k], rots[k])
        worst = max(worst, float((a - bx.outputs[k]).abs().max()) / max(float(a.abs().max()), 1e-30))
        va, wa = tt.insert_central_slices_rfft_3d(slices[k], shapes[k], rots[k])
        vb, wb = bi.outputs[k]
        worst = max(worst, float(torch.linalg.vector_norm(va - vb) / torch.linalg.vector_norm(va)),
                    float(torch.linalg.vector_norm(wa - wb) / torch.linalg.vector_norm(wa)))
    n_rot = int(counts.sum())
    line = {
        "metric": "slices/s batched extract + insert over 32 volumes of mixed box size (graph replay)",
        "value": 2 * n_rot / (ms_b / 1e3), "unit": "slices/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_b, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64",
        "data": "synthetic (seeded blob volumes, uniform SO(3) rotations, N(0,1) c64 slices)",
        "config": {"workload": f"32 volumes, boxes {sorted(set(int(v) for v in sizes))}, {n_rot} rotations, "
                               "extract + insert", "parallelism": "single GPU"},
        "one_call_loop_ms": ms_l, "speedup_vs_loop": ms_l / ms_b,
        "tiles": {"extract": sorted(set(bx.tile_shapes)), "insert": sorted(set(bi.tile_shapes))},
        "parity_vs_one_call": {"worst_relative": worst, "bar": "<= 1e-5"},
    }
    bx.close()
    bi.close()
    print(json.dumps(line), flush=True)


def torch_project(volume, R, pad):
    """The reference's projection (project.py:44-93) composed from torch
    operations, with the plain-torch gather above in chunks of 64 poses."""
    import torch.nn.functional as Fn

    d = volume.shape[-1]
    vol = Fn.pad(volume, [pad] * 6)
    n = vol.shape[-1]
    f = torch.fft.fftshift(torch.fft.fftfreq(n, device=vol.device))
    r = torch.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2)
    vol = vol * torch.special.sinc(r) ** 2
    dft = torch.fft.fftshift(vol, dim=(-3, -2, -1))
    dft = torch.fft.fftshift(torch.fft.rfftn(dft, dim=(-3, -2, -1)), dim=(-3, -2))
    sl = torch.cat([torch_extract(dft, (n, n, n), R[k:k + 64]) for k in range(0, R.shape[0], 64)])
    sl = torch.fft.ifftshift(sl, dim=(-2,))
    proj = torch.fft.ifftshift(torch.fft.irfftn(sl, dim=(-2, -1)), dim=(-2, -1))
    return proj[..., pad:pad + d, pad:pad + d]


def run_project(args, device):
    """project_3d_to_2d and backproject_2d_to_3d end to end at 128^3
    (blob volume, pad_factor 2 -> 256^3, 512 poses): the fused
    pre/post kernels, torch's FFTs and the brick gather / scatter; the
    projection also against its torch composition on the same GPU."""
    import teamtomo_amd as tt
    from teamtomo_amd.synthetic import blob_volume, uniform_rotations

    vol = blob_volume(128, 200, seed=2, device=device)
    R = uniform_rotations(512, seed=0, device=device)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.steps

    with torch.no_grad():
        proj = tt.project_3d_to_2d(vol, R, pad_factor=2.0)
        ms_p = timed(lambda: tt.project_3d_to_2d(vol, R, pad_factor=2.0))
        ms_b = timed(lambda: tt.backproject_2d_to_3d(proj, R, pad_factor=2.0))
        ref = torch_project(vol, R, 64)
        ms_t = timed(lambda: torch_project(vol, R, 64))
    rel = float(torch.linalg.vector_norm(proj - ref) / torch.linalg.vector_norm(ref))
    line = {
        "metric": "poses/s project_3d_to_2d at 128^3 (pad 2 -> 256^3), 512 poses",
        "value": 512 / (ms_p / 1e3), "unit": "poses/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_p, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/c64",
        "data": "synthetic (seeded 200-blob 128^3 volume, uniform SO(3) rotations)",
        "config": {"workload": "project_3d_to_2d 128^3 pad_factor 2, 512 poses; backproject_2d_to_3d of the result",
                   "parallelism": "single GPU"},
        "backproject_ms": ms_b,
        "torch_same_gpu": {"project_ms": ms_t, "speedup": ms_t / ms_p, "rel_l2_vs_ours": rel,
                           "note": "torch composition with the plain advanced-indexing gather"},
    }
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.write_traffic:
        write_traffic(args.write_traffic)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md, not launched)"}))
        sys.exit(2)
    device = torch.device("cuda:0")
    torch.cuda.set_device(device)
    if args.copy_probe:
        n = args.copy_probe << 18  # float32 elements
        src = torch.ones(n, device=device)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": 4 * n, "copies": 5}), flush=True)
        return
    if args.workload == "lines2d":
        run_lines2d(args, device)
        return
    if args.workload == "insert3d":
        run_insert3d(args, device)
        return
    if args.workload == "batch":
        run_batch(args, device)
        return
    if args.workload == "project":
        run_project(args, device)
        return
    from teamtomo_amd import _lib, set_gather_path

    _lib.load()
    set_gather_path(args.gather)
    V, R, gS = make_inputs(device)
    step = Step(V, R, gS, args.deterministic)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(args.steps):
        step(timed=True)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = start.elapsed_time(stop)
    ms_step = total_ms / args.steps
    ext_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in step.events]))
    adj_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in step.events]))
    ext_b, adj_b = algorithmic_bytes()
    kernels = {
        "extract": {"ms": ext_ms, "bytes": ext_b, "GBs": ext_b / ext_ms / 1e6,
                    "launch": "tt_extract_c64: plane_rec_kernel + gather_kernel (brick units and zero rows)"},
        "adjoint": {"ms": adj_ms, "bytes": adj_b, "GBs": adj_b / adj_ms / 1e6,
                    "launch": "tt_extract_bwd_c64 volume gradient: fill_zero_kernel + band_max_kernel "
                              "(|f|-band maxima of every 64th slice) + plane_rec_kernel + brick_kernel (adjoint) + "
                              "split_combine_kernel + halo_combine_kernel, then the guarded exact-scale redo "
                              "(maxabs_chunks_kernel, brick_kernel, split_combine_kernel, halo_combine_kernel; "
                              "each exits at once unless a value exceeded its brick's sampled limit)"},
    }
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    achieved = kernels[dom]["GBs"]
    step_bytes = ext_b + adj_b
    step_gbs = step_bytes / ms_step / 1e6
    value = N_ROT / (ms_step / 1e3)

    traffic, traffic_note = measured_traffic(dom)

    line = {
        "metric": "slices/s fwd+adjoint at 256^3 c64 vs plain torch on same GPU; % HBM roofline",
        "value": value,
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded 200-blob 256^3 volume, uniform SO(3) rotations, N(0,1) c64 gradients)",
        "config": {"workload": "extract+adjoint 256^3 rfft (256,256,129) c64, 4096 rotations",
                   "volume": [N, N, N // 2 + 1], "rotations": N_ROT, "parallelism": "single GPU",
                   "deterministic": bool(args.deterministic)},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_per_algorithmic_byte": traffic / kernels[dom]["bytes"] if traffic else None,
                     "traffic_note": traffic_note},
        "roofline_step": {"achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": step_gbs / HBM_PEAK_GBS, "bytes": step_bytes},
        "kernels": kernels,
        "wall_s": wall,
        "copy_GBs": copy_bandwidth(device),
    }
    if not args.no_torch:
        n_t = N_ROT if args.torch else 512
        try:
            tv, tms, chunk = time_torch_comparator(V, R, gS, n_rot=n_t)
            line["torch_same_gpu"] = {
                "value": tv, "unit": "slices/s", "ms_per_sample": tms, "rotations": n_t, "rotation_chunk": chunk,
                "speedup": value / tv,
                "sample": f"extract + adjoint of the first {n_t} of the {N_ROT} rotations at {N}^3, "
                          f"chunks of {chunk}, reported per slice"}
        except Exception as exc:  # report, do not hide the main result
            line["torch_same_gpu"] = {"error": repr(exc)[:300]}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(V, R, gS, step)
        line["cpu_baseline"]["torch_cpu"] = torch_cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
