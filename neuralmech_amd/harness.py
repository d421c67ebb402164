"""Throughput harnesses of ``projects/18_simple_wave_fd`` over this library's kernels.

* :func:`roofline` -- ``2D_wave_roofline.py``: the copy ceiling (1 GiB device copy), the fp32
  FMA peak (``wfd_fma_peak``) and every step kernel of the reference layout on a 1600 x 3200 grid
  for 2000 steps, reported as Gflop/s at the reference's 18 flop per dof and its intensity
  ``18 / (3 * bytes per entry)`` (``2D_wave_roofline.py:31-36,105-117``);
* :func:`scaling` -- ``2D_wave_scaling.py``: dofs against seconds per step over grid sizes
  ``unique(logspace(0.5, 5, 56))``, steps replayed from a captured graph of 20
  (``2D_wave_scaling.py:24-31,77-99``);
* :func:`save_csv` -- the space-separated columns the reference's plotting reads.

Every launch goes to torch's current stream; graph capture records the same launches.
"""

from __future__ import annotations

import ctypes
import math
import time
from pathlib import Path

import numpy as np
import torch

from . import _lib, reference_steps as rs
from ._lib import check

FLOPS_PER_DOF = 18


def save_csv(path, **columns) -> None:
    """Write equal-length columns, space separated, one header line with the column names."""
    names = list(columns)
    cols = [list(np.asarray(columns[n]).ravel()) for n in names]
    if len({len(c) for c in cols}) > 1:
        raise ValueError("columns differ in length")
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w") as f:
        f.write(" ".join(names) + "\n")
        for row in zip(*cols):
            f.write(" ".join(str(v) for v in row) + "\n")


def _events():
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def copy_ceiling(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Device-to-device copy rate in bytes/s (read + write), as ``2D_wave_roofline.py:46-57``."""
    a = torch.zeros(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = _events()
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    return reps * 2 * a.numel() * 4 / (e0.elapsed_time(e1) / 1e3)


def fma_peak(blocks: int = 4096, n: int = 4096) -> float:
    """fp32 FMA rate in flop/s from the register FMA chain (``2D_wave_roofline.py:59-74``)."""
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    lib = _lib.load()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib.wfd_fma_peak(out.data_ptr(), blocks, n, st))
    e0, e1 = _events()
    e0.record()
    check(lib.wfd_fma_peak(out.data_ptr(), blocks, n, st))
    e1.record()
    torch.cuda.synchronize()
    return blocks * 256 * n * 8 / (e0.elapsed_time(e1) / 1e3)


def _graph_seconds_per_step(body, steps_per_call: int, steps: int, chunk_calls: int = 20) -> float:
    """Capture ``chunk_calls`` calls of ``body`` in a graph, replay it to cover ``steps`` steps;
    seconds per step."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(chunk_calls):  # warm-up outside the graph
            body()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(chunk_calls):
            body()
    g.replay()  # the first replay uploads the graph
    torch.cuda.synchronize()
    reps = max(1, steps // (chunk_calls * steps_per_call))
    e0, e1 = _events()
    e0.record()
    for _ in range(reps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / (reps * chunk_calls * steps_per_call)


def _problem(nx, ny, lengths=(1.0, 2.0), wavespeed=1.0, safety=0.95):
    dx, dy = lengths[0] / (nx - 1), lengths[1] / (ny - 1)
    dt = safety * min(dx, dy) / wavespeed / math.sqrt(2)
    return dx, dy, dt


def _kernel_body(kind: str, nx: int, ny: int, dx, dy, dt, c=1.0):
    """(body, steps per call, bytes per entry) of one reference-layout kernel, fields filled
    with random values as in the reference harness."""
    gen = torch.Generator(device="cuda").manual_seed(0)
    if kind in ("basic", "padded", "shared", "temporal", "scaled"):
        mp = ny if kind == "basic" else ((ny + 31) // 32) * 32
        if kind == "scaled":
            mp = ((ny + 2 + 127) // 128) * 128
    else:  # mixed
        mp = ny + 2 + (ny % 2)
    if kind == "mixed":
        U = torch.rand((2, nx, mp), generator=gen, device="cuda").to(torch.float16)
        f = [U[0], U[1]]

        def body():
            rs.mixed_step(f[0], f[1], f[0], c, dt, dx, dy, nx, ny, mp)
            f[0], f[1] = f[1], f[0]

        return body, 1, 2
    if kind == "scaled":
        q = [torch.zeros((nx, mp), dtype=torch.int16, device="cuda") for _ in range(2)]
        sc = [torch.zeros((nx, mp // 32), dtype=torch.float32, device="cuda") for _ in range(2)]

        def body():
            rs.scaled_step(q[0], sc[0], q[1], sc[1], q[0], sc[0], c, dt, dx, dy, nx, ny, mp)
            q[0], q[1] = q[1], q[0]
            sc[0], sc[1] = sc[1], sc[0]

        return body, 1, 2 + 4 / 32
    U = torch.rand((4 if kind == "temporal" else 2, nx, mp), generator=gen, device="cuda")
    f = list(U)
    if kind == "temporal":
        def body():
            rs.temporal_step(f[0], f[1], f[2], f[3], c, dt, dx, dy, nx, ny, mp, 8)
            f[0], f[1], f[2], f[3] = f[2], f[3], f[0], f[1]

        return body, 8, 4
    step = {"basic": lambda a, b: rs.basic_step(a, b, a, c, dt, dx, dy, nx, ny),
            "padded": lambda a, b: rs.padded_step(a, b, a, c, dt, dx, dy, nx, ny, mp),
            "shared": lambda a, b: rs.shared_step(a, b, a, c, dt, dx, dy, nx, ny, mp)}[kind]

    def body():
        step(f[0], f[1])
        f[0], f[1] = f[1], f[0]

    return body, 1, 4


KERNELS = ("basic", "padded", "shared", "temporal", "mixed", "scaled")


def roofline(nx: int = 1600, ny: int = 3200, steps: int = 2000, kernels=KERNELS) -> dict:
    """Ceilings and every kernel's Gflop/s at 18 flop per dof (the reference's problem)."""
    dx, dy, dt = _problem(nx, ny)
    dofs = (nx - 2) * (ny - 2)
    out = {"bandwidth_GBs": copy_ceiling() / 1e9, "peak_Gflops": fma_peak() / 1e9, "grid": [nx, ny],
           "steps": steps, "kernels": []}
    for kind in kernels:
        body, per_call, nbytes = _kernel_body(kind, nx, ny, dx, dy, dt)
        sec = _graph_seconds_per_step(body, per_call, steps, chunk_calls=20 if per_call == 1 else 4)
        out["kernels"].append({"kernel": kind, "intensity": FLOPS_PER_DOF / (3 * nbytes),
                               "gflops": FLOPS_PER_DOF * dofs / sec / 1e9, "gdofs_per_s": dofs / sec / 1e9,
                               "s_per_step": sec})
        torch.cuda.empty_cache()
    return out


def scaling(kind: str = "padded", max_resolution: int = 100000, steps: int = 2000, memory: float = 0.9):
    """``(dofs, seconds per step)`` over square grids ``unique(logspace(0.5, 5, 56))`` up to
    ``max_resolution``, stopping where the two fields exceed ``memory`` of the free memory."""
    dofs, timings = [], []
    for res in np.unique(np.logspace(0.5, 5.0, 56).astype(np.int64)):
        n = int(res)
        if n < 3 or n > max_resolution:
            continue
        itemsize = 2 if kind == "mixed" else 4
        free, _ = torch.cuda.mem_get_info()
        if 2 * n * (((n + 31) // 32) * 32 + 2) * itemsize > memory * free:
            break
        dx, dy, dt = _problem(n, n, lengths=(1.0, 1.0))
        body, per_call, _ = _kernel_body(kind, n, n, dx, dy, dt)
        timings.append(_graph_seconds_per_step(body, per_call, steps))
        dofs.append((n - 2) * (n - 2))
        del body
        torch.cuda.empty_cache()
    return dofs, timings


def timed(fn, *args, **kw):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn(*args, **kw)
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0
