"""Host-side solver objects over the wavefd C ABI: grid layout, medium, point lists, forward runs
and the checkpointed gradient plan.  Buffers are torch tensors on the ROCm device; the library
borrows their pointers and launches on torch's current stream.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import WFD_F32, WFD_F64, check

_DTYPES = {torch.float32: WFD_F32, torch.float64: WFD_F64}


_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _stream_ptr(device: torch.device):
    """torch's current stream on ``device`` as a ``hipStream_t`` (the raw handle query skips the
    construction of a ``torch.cuda.Stream`` object: a few microseconds per launching call)."""
    if device.type != "cuda":
        raise RuntimeError("the wave kernels run on a ROCm GPU; got a tensor on " + str(device))
    if _RAW_STREAM is not None:
        return ctypes.c_void_p(_RAW_STREAM(device.index if device.index is not None else torch.cuda.current_device()))
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stable_dt(spacing: Sequence[float], vmax: float, order: int = 2) -> float:
    """Largest stable leapfrog time step for the given spacing, maximum speed and order.

    ``vmax dt <= 2 / sqrt(sum_d S / h_d^2)``, ``S = |a_0| + 2 sum |a_k|``.  For order 2 in 2D this
    is the reference's ``min(dx, dy) / c / sqrt(2)`` (``2D_wave_cupy.py:38``); the cuwave call it
    replaces is ``stable_dt(dx, wavespeed, space_order)`` (``projects/18_fwi/helper.py:52``).
    """
    lib = _lib.load()
    h = (ctypes.c_double * len(spacing))(*spacing)
    dt = lib.wfd_stable_dt(len(spacing), h, float(vmax), int(order))
    if not dt > 0:
        raise ValueError(_lib.last_error())
    return dt


class WaveGrid:
    """A node grid (boundary ring included), its spacing, time step, order, dtype and batch size.

    Fields live in the halo layout of ``include/wavefd.h``: a tensor ``(nshots, elems)``;
    :meth:`nodes` gives the ``(nshots, *shape)`` strided view of the nodes.
    """

    def __init__(self, shape, spacing, dt, order=2, dtype=torch.float32, nshots=1, device=None):
        shape = tuple(int(s) for s in shape)
        spacing = tuple(float(h) for h in (spacing if np.ndim(spacing) else [spacing] * len(shape)))
        if len(spacing) != len(shape):
            raise ValueError("spacing must have one entry per axis")
        if dtype not in _DTYPES:
            raise TypeError("dtype must be torch.float32 or torch.float64")
        self.shape, self.spacing, self.dt, self.order = shape, spacing, float(dt), int(order)
        self.dtype, self.nshots = dtype, int(nshots)
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        lib = _lib.load()
        self.c = _lib.Grid()
        n = (ctypes.c_int64 * 3)(*shape, *([1] * (3 - len(shape))))
        h = (ctypes.c_double * 3)(*spacing, *([1.0] * (3 - len(spacing))))
        check(lib.wfd_grid_init(ctypes.byref(self.c), len(shape), n, h, self.dt, self.order,
                                _DTYPES[dtype], self.nshots))
        self.ndim = len(shape)
        self.elems, self.offset = int(self.c.elems), int(self.c.offset)
        self.ld, self.plane = int(self.c.ld), int(self.c.plane)
        self.nrows = int(lib.wfd_points_nrows(ctypes.byref(self.c)))

    # ------------------------------------------------------------------ layout helpers
    @property
    def interior_cells(self) -> int:
        return math.prod(n - 2 for n in self.shape)

    def zeros(self, nshots=None, dtype=None, device=None) -> torch.Tensor:
        return torch.zeros((self.nshots if nshots is None else nshots, self.elems),
                           dtype=dtype or self.dtype, device=device or self.device)

    def nodes(self, field: torch.Tensor) -> torch.Tensor:
        """Strided ``(nshots, *shape)`` view of the nodes of a layout tensor ``(nshots, elems)``."""
        b = field.shape[0]
        o = self.offset
        if self.ndim == 1:
            return field[:, o:o + self.ld][:, : self.shape[0]]
        if self.ndim == 2:
            n0, n1 = self.shape
            return field[:, o:o + n0 * self.ld].view(b, n0, self.ld)[:, :, :n1]
        n0, n1, n2 = self.shape
        return field[:, o:o + n0 * self.plane * self.ld].view(b, n0, self.plane, self.ld)[:, :, :n1, :n2]

    def from_nodes(self, arr, nshots=None, dtype=None) -> torch.Tensor:
        """Layout tensor holding ``arr`` (``shape`` or ``(b, *shape)``) at the nodes, zero elsewhere."""
        t = torch.as_tensor(arr)
        if t.dim() == self.ndim:
            t = t.unsqueeze(0)
        if tuple(t.shape[1:]) != self.shape:
            raise ValueError(f"expected node array of shape {self.shape}, got {tuple(t.shape[1:])}")
        b = t.shape[0] if nshots is None else nshots
        out = self.zeros(nshots=b, dtype=dtype)
        self.nodes(out).copy_(t.expand(b, *self.shape))
        return out

    def flat_nodes(self, points) -> np.ndarray:
        """Flat C-order node indices of ``points`` (index tuples or flat ints)."""
        a = np.asarray(points, dtype=np.int64)
        if a.size == 0:
            return np.zeros(0, np.int64)
        if a.ndim == 2:
            return np.ravel_multi_index(tuple(a.T), self.shape).astype(np.int64)
        return a.reshape(-1)

    # ------------------------------------------------------------------ launch shapes
    def set_launch(self, variant: int, rows: int) -> None:
        """Force the kernel variant and rows (2D) / planes (3D) per wave for this grid shape;
        ``rows = 0`` returns to the tuned or default shape."""
        check(_lib.load().wfd_set_config(ctypes.byref(self.c), int(variant), int(rows)))

    def launch(self):
        v, r = ctypes.c_int32(), ctypes.c_int32()
        check(_lib.load().wfd_get_config(ctypes.byref(self.c), ctypes.byref(v), ctypes.byref(r)))
        return v.value, r.value

    def tune(self, medium: "Medium", steps: int = 20) -> dict:
        """Time every launch shape on scratch fields and keep the fastest for this grid shape."""
        a, b = self.zeros(), self.zeros()
        v, r, us = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_double()
        check(_lib.load().wfd_tune(ctypes.byref(self.c), a.data_ptr(), b.data_ptr(), ctypes.byref(medium.c),
                                   int(steps), ctypes.byref(v), ctypes.byref(r), ctypes.byref(us),
                                   _stream_ptr(self.device)))
        return {"variant": v.value, "rows": r.value, "us_per_step": us.value}

    def stencil_weights(self):
        """``(wc, w[ndim, K])`` as the kernels use them (rounded to the grid dtype)."""
        K = self.order // 2
        wc = ctypes.c_double()
        w = (ctypes.c_double * 12)()
        check(_lib.load().wfd_stencil_weights(ctypes.byref(self.c), ctypes.byref(wc), w))
        return wc.value, np.array(w[: self.ndim * K]).reshape(self.ndim, K)


class PointSet:
    """Sources or receivers: per shot a list of interior nodes, stored as a device CSR.

    ``nodes`` is one list of points (index tuples or flat node indices) shared by every shot, or,
    with ``per_shot=True``, one such list per shot (ragged allowed).
    """

    def __init__(self, grid: WaveGrid, nodes, per_shot: bool = False, device=None):
        self.grid = grid
        per_shot = _per_shot(nodes, grid.nshots, per_shot)
        flat = [grid.flat_nodes(p) for p in per_shot]
        self.counts = [len(f) for f in flat]
        self.count = max(self.counts) if self.counts else 0
        lib = _lib.load()
        nr = grid.nrows
        rp = np.zeros((grid.nshots, nr + 1), np.int32)
        col = np.zeros((grid.nshots, max(self.count, 1)), np.int32)
        ids = np.zeros((grid.nshots, max(self.count, 1)), np.int32)
        rows = np.zeros((grid.nshots, max(self.count, 1)), np.int32)
        i32 = ctypes.POINTER(ctypes.c_int32)
        for s, f in enumerate(flat):
            f = np.ascontiguousarray(f, np.int64)
            r = np.zeros(nr + 1, np.int32)
            c = np.zeros(max(len(f), 1), np.int32)
            i = np.zeros(max(len(f), 1), np.int32)
            w = np.zeros(max(len(f), 1), np.int32)
            check(lib.wfd_points_csr(ctypes.byref(grid.c), len(f), f.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     r.ctypes.data_as(i32), c.ctypes.data_as(i32), i.ctypes.data_as(i32),
                                     w.ctypes.data_as(i32)))
            rp[s] = r
            col[s, : len(f)] = c[: len(f)]
            ids[s, : len(f)] = i[: len(f)]
            rows[s, : len(f)] = w[: len(f)]
        dev = device or grid.device
        self.nodes_flat = flat
        self.row_ptr = torch.from_numpy(rp).to(dev)
        self.col = torch.from_numpy(col).to(dev)
        self.id = torch.from_numpy(ids).to(dev)
        self.row = torch.from_numpy(rows).to(dev)

    def struct(self, data: Optional[torch.Tensor]) -> Optional[_lib.Points]:
        """C descriptor with ``data`` ``(nshots, steps, width)`` as signal or trace buffer."""
        if self.count == 0 or data is None:
            return None
        if data.dim() != 3 or data.shape[0] != self.grid.nshots or not data.is_contiguous():
            raise ValueError("point data must be a contiguous (nshots, steps, width) tensor")
        if data.shape[2] < self.count or data.dtype != self.grid.dtype:
            raise ValueError("point data too narrow or of the wrong dtype")
        p = _lib.Points()
        p.count, p.nrows = self.count, self.grid.nrows
        p.row_ptr, p.col, p.id = self.row_ptr.data_ptr(), self.col.data_ptr(), self.id.data_ptr()
        p.row = self.row.data_ptr()
        p.data = data.data_ptr()
        p.width = data.shape[2]
        p.nsteps = data.shape[1]
        p.shot_stride = data.shape[1] * data.shape[2]
        return p

    def gather(self, field: torch.Tensor) -> torch.Tensor:
        """Values of a layout field at the points, ``(nshots, count)`` (zero-padded)."""
        nodes = self.grid.nodes(field).reshape(field.shape[0], -1)
        out = torch.zeros(field.shape[0], max(self.count, 1), dtype=field.dtype, device=field.device)
        for s, f in enumerate(self.nodes_flat):
            if len(f):
                out[s, : len(f)] = nodes[s, torch.as_tensor(f, device=field.device)]
        return out[:, : self.count]


def _per_shot(nodes, nshots, per_shot):
    if nodes is None:
        return [[] for _ in range(nshots)]
    if per_shot:
        nodes = list(nodes)
        if len(nodes) != nshots:
            raise ValueError(f"per-shot point lists: expected {nshots}, got {len(nodes)}")
        return nodes
    return [nodes] * nshots


class Medium:
    """``kc = v^2 dt^2`` per node (or a constant) and optional sponge profiles.

    ``velocity`` is a node array of the grid shape (shared by all shots) or ``(nshots, *shape)``;
    ``beta_profiles`` one damping profile per axis (slowest first), ``beta = d dt / 2``.  On a GPU
    the values are checked once here (``check``): NaN/Inf or ``kc <= 0`` raise ``WaveError``
    (status WFD_ENONFINITE), a ``kc`` above the CFL limit of the grid's order and spacing raises
    it with status WFD_EINVAL.
    """

    def __init__(self, grid: WaveGrid, velocity=None, kc=None, kconst=None, beta_profiles=None, check=True):
        self.grid = grid
        self.coef = None
        self.kconst = 0.0
        if velocity is not None or kc is not None:
            if kc is None:
                v = torch.as_tensor(velocity, dtype=torch.float64)
                kc = (v * grid.dt) ** 2
            kc = torch.as_tensor(kc).to(device=grid.device, dtype=torch.float64)
            nb = 1 if kc.dim() == grid.ndim else kc.shape[0]
            if nb not in (1, grid.nshots):
                raise ValueError("per-shot velocity needs nshots entries")
            self.coef = grid.from_nodes(kc, nshots=nb, dtype=grid.dtype)
            self.coef_shot_stride = 0 if nb == 1 else grid.elems
        elif kconst is not None:
            self.kconst = float(kconst)
            self.coef_shot_stride = 0
        else:
            raise ValueError("give velocity, kc or kconst")
        self.beta = None
        if beta_profiles is not None:
            if len(beta_profiles) != grid.ndim:
                raise ValueError("one damping profile per axis")
            parts = []
            for n, p in zip(grid.shape, beta_profiles):
                p = torch.as_tensor(p, dtype=torch.float64).reshape(-1)
                if p.numel() != n:
                    raise ValueError("damping profile length must equal the axis node count")
                parts.append(p)
            self.beta = torch.cat(parts).to(device=grid.device, dtype=grid.dtype)
        self.c = _lib.Medium()
        self.c.coef = None if self.coef is None else self.coef.data_ptr()
        self.c.kconst = self.kconst
        self.c.coef_shot_stride = self.coef_shot_stride
        self.c.beta = None if self.beta is None else self.beta.data_ptr()
        if check and grid.device.type == "cuda":
            self.validate()

    def kc_nodes(self) -> torch.Tensor:
        if self.coef is None:
            return torch.full((1, *self.grid.shape), self.kconst, dtype=torch.float64, device=self.grid.device)
        return self.grid.nodes(self.coef)

    def validate(self):
        """Raise ``WaveError`` on NaN/Inf or non-positive coefficients (status WFD_ENONFINITE) or on a
        coefficient above the CFL limit (WFD_EINVAL); host-synchronising."""
        check(_lib.load().wfd_check_medium(ctypes.byref(self.grid.c), ctypes.byref(self.c),
                                           _stream_ptr(self.grid.device)))


def _sig_tensor(grid: WaveGrid, pts: Optional[PointSet], sig, nsteps: int):
    """Signal as a contiguous ``(nshots, nsteps, count)`` tensor of the grid dtype."""
    if pts is None or pts.count == 0:
        return None
    if sig is None:
        raise ValueError("sources need a signal")
    s = torch.as_tensor(sig, dtype=grid.dtype).to(grid.device)
    if s.dim() == 1:
        s = s[None, :, None].expand(grid.nshots, -1, pts.count)
    elif s.dim() == 2:
        s = s[None].expand(grid.nshots, -1, -1)
    if s.shape[1] < nsteps:
        raise ValueError(f"signal has {s.shape[1]} steps, the run needs {nsteps}")
    if s.shape[2] < pts.count:
        raise ValueError("signal has fewer columns than sources")
    return s.contiguous()


class WaveState:
    """The two newest time levels of a run and the step they stand at: ``prev = u^{n-1}``,
    ``cur = u^n`` (layout tensors ``(nshots, elems)``), ``step = n``.

    :func:`run` and :class:`ForwardRun` advance it in place -- the levels and the step -- so the
    same object passed again continues the run whatever the parity of the step counts (the
    library rotates two buffers in place, ``2D_wave_cuda_padding.py:83-85``, so after an odd
    count the newest level sits in the other buffer).  It reads like the pair ``(prev, cur)``.
    """

    __slots__ = ("prev", "cur", "step")

    def __init__(self, prev: torch.Tensor, cur: torch.Tensor, step: int = 0):
        if prev is cur or prev.data_ptr() == cur.data_ptr():
            raise ValueError("prev and cur must be distinct fields")
        self.prev, self.cur, self.step = prev, cur, int(step)

    @classmethod
    def rest(cls, grid: WaveGrid, step: int = 0) -> "WaveState":
        """Zero fields (the medium at rest) standing at ``step``."""
        return cls(grid.zeros(), grid.zeros(), step)

    def clone(self) -> "WaveState":
        return WaveState(self.prev.clone(), self.cur.clone(), self.step)

    def __iter__(self):
        return iter((self.prev, self.cur))

    def __getitem__(self, i):
        return (self.prev, self.cur)[i]

    def __len__(self):
        return 2


def _as_state(grid: WaveGrid, state, step0) -> WaveState:
    if state is None:
        return WaveState.rest(grid, 0 if step0 is None else step0)
    if not isinstance(state, WaveState):
        raise TypeError("state must be a solver.WaveState: a plain (u_prev, u_cur) pair cannot record "
                        "which buffer holds the newest level after an odd number of steps")
    if step0 is not None and int(step0) != state.step:
        raise ValueError(f"step0 = {step0} but the state stands at step {state.step}")
    return state


class ForwardRun:
    """A forward run staged once: the signal, a trace buffer with a row for every step up to
    ``total`` and the C descriptors are built at construction, so :meth:`advance` only launches
    (``wfd_forward``; a replay of a ``wfd_fgraph`` captured once per buffer orientation with
    ``graph_chunk``; ``wfd_forward2`` with ``two_step``).

    ``traces[s, n, r] = u^{n+1}`` at receiver ``r`` of shot ``s`` for the absolute step ``n``;
    ``state`` (a :class:`WaveState`, rest at step 0 when omitted) is advanced in place.
    ``graph_chunk`` > 0 (even) replays a hipGraph of that many steps -- the reference's captured
    20-step graph, ``2D_wave_cuda_graph.py:73-79,91-96`` -- and runs the remainder as plain
    launches; ``two_step`` > 0 (2D) runs two steps per launch with that many output rows per wave
    (2, 4 or 8: the intermediate level in LDS; 16 or 32: in registers) on two work fields.  Every
    form gives the same bits.
    """

    def __init__(self, grid: WaveGrid, medium: Medium, total: int, src: Optional[PointSet] = None, sig=None,
                 rec: Optional[PointSet] = None, state: Optional[WaveState] = None, graph_chunk: int = 0,
                 two_step: int = 0):
        self.grid, self.medium, self.total = grid, medium, int(total)
        self.state = _as_state(grid, state, None)
        if self.state.step > self.total:
            raise ValueError(f"the state stands at step {self.state.step}, past total = {self.total}")
        self.graph_chunk, self.two_step = int(graph_chunk), int(two_step)
        if self.graph_chunk and (self.graph_chunk < 2 or self.graph_chunk % 2):
            raise ValueError("graph_chunk must be even and >= 2")
        self.src, self.rec = src, rec
        self.sig = _sig_tensor(grid, src, sig, self.total)
        self.traces = None
        if rec is not None and rec.count:
            self.traces = torch.zeros(grid.nshots, self.total, rec.count, dtype=grid.dtype, device=grid.device)
        self._sp = src.struct(self.sig) if src is not None else None
        self._rp = rec.struct(self.traces) if rec is not None else None
        self._work = (grid.zeros(), grid.zeros()) if self.two_step else None
        self._graphs = {}
        # the call's fixed arguments, built once: advance() only adds the buffers, step and count
        self._lib = _lib.load()
        sp, rp = self._pts()
        self._fixed = (ctypes.byref(grid.c), ctypes.byref(medium.c), sp, rp)

    def _pts(self):
        return (ctypes.byref(self._sp) if self._sp else None, ctypes.byref(self._rp) if self._rp else None)

    def prepare(self) -> None:
        """Capture the graph for the current buffer orientation now (outside any timing)."""
        if self.graph_chunk:
            self._graph()

    def _graph(self):
        key = (self.state.prev.data_ptr(), self.state.cur.data_ptr())
        gr = self._graphs.get(key)
        if gr is None:
            gr = ForwardGraph(self.grid, self.state.prev, self.state.cur, self.medium, self._sp, self._rp,
                              self.graph_chunk)
            self._graphs[key] = gr
        return gr

    def advance(self, nsteps: int) -> WaveState:
        """Run ``nsteps`` more steps on the stream; the state moves to ``step + nsteps``."""
        nsteps = int(nsteps)
        st, g = self.state, self.grid
        if nsteps < 0 or st.step + nsteps > self.total:
            raise ValueError(f"{nsteps} steps from step {st.step} pass total = {self.total}")
        if nsteps == 0:
            return st
        gp, mp, sp, rp = self._fixed
        stream = _stream_ptr(g.device)
        if self.two_step:
            wc, wd = self._work
            check(self._lib.wfd_forward2(gp, st.prev.data_ptr(), st.cur.data_ptr(), wc.data_ptr(), wd.data_ptr(), mp,
                                         sp, rp, st.step, nsteps, self.two_step, stream))
            st.step += nsteps  # (u^{N-1}, u^N) stay in (prev, cur)
            return st
        if self.graph_chunk:
            self._graph().run(st.step, nsteps)
        else:
            check(self._lib.wfd_forward(gp, st.prev.data_ptr(), st.cur.data_ptr(), mp, sp, rp, st.step, nsteps,
                                        stream))
        if nsteps % 2:  # the newest level is now in the buffer that held u^{n-1}
            st.prev, st.cur = st.cur, st.prev
        st.step += nsteps
        return st


class ForwardGraph:
    """``wfd_fgraph``: ``chunk`` steps of the forward run on the fields ``(u_a, u_b)`` captured once
    into a hipGraph with its own device step counter (the reference's capture-once, replay-many
    graph, ``2D_wave_cuda_graph.py:73-79,91-96``).  The buffers, medium and point descriptors it
    names must stay alive while it exists."""

    def __init__(self, grid: WaveGrid, u_a, u_b, medium: Medium, src_struct, rec_struct, chunk: int):
        self.grid = grid
        self._keep = (u_a, u_b, medium, src_struct, rec_struct)
        self._p = ctypes.c_void_p()
        check(_lib.load().wfd_fgraph_create(ctypes.byref(self._p), ctypes.byref(grid.c), _ptr(u_a), _ptr(u_b),
                                            ctypes.byref(medium.c),
                                            ctypes.byref(src_struct) if src_struct else None,
                                            ctypes.byref(rec_struct) if rec_struct else None, int(chunk)))

    def run(self, step0: int, nsteps: int) -> None:
        check(_lib.load().wfd_fgraph_run(self._p, int(step0), int(nsteps), _stream_ptr(self.grid.device)))

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            try:
                _lib.load().wfd_fgraph_destroy(p)
            except Exception:
                pass
            self._p = None


def run(grid: WaveGrid, medium: Medium, nsteps: int, src: Optional[PointSet] = None, sig=None,
        rec: Optional[PointSet] = None, state: Optional[WaveState] = None, step0: Optional[int] = None,
        graph_chunk: int = 0, two_step: int = 0):
    """Advance ``nsteps`` steps; returns ``(traces, state)``.

    ``traces[s, n, r] = u^{step0+n+1}`` at receiver ``r`` of shot ``s``.  ``state`` is a
    :class:`WaveState`: omitted, the run starts at rest at ``step0`` (default 0); given, it
    continues from ``state.step`` and is advanced in place (``step0``, if passed, must equal it),
    so ``run(N1)`` followed by ``run(N2, state=...)`` gives the bits of ``run(N1 + N2)``.
    ``graph_chunk`` and ``two_step`` select the launch form as in :class:`ForwardRun`.
    """
    st = _as_state(grid, state, step0)
    s0 = st.step
    fr = ForwardRun(grid, medium, s0 + int(nsteps), src, sig, rec, st, graph_chunk, two_step)
    fr.advance(nsteps)
    out = None if fr.traces is None else fr.traces[:, s0:]
    return out, st


def kc_gradient(grid: WaveGrid, Gn: torch.Tensor, mu_src, sig, src: Optional[PointSet], kc) -> torch.Tensor:
    """``dJ/dkc`` on the nodes from the imaging sum ``Gn = sum_n mu^{n+1} L u^n`` (node array,
    modified in place): the source term ``sum_n mu^{n+1}(x_s) sig^n_s`` is added at the source
    nodes (``mu_src`` ``(nshots, nsteps, S)``, row n = mu^{n+1}), then ``-G / kc`` on the interior."""
    if mu_src is not None:
        contrib = (mu_src.double() * sig.double()[:, :, : src.count]).sum(1)  # (nshots, count)
        flat = Gn.reshape(-1)
        for s, f in enumerate(src.nodes_flat):
            if len(f):
                flat.index_add_(0, torch.as_tensor(f, device=Gn.device), contrib[s, : len(f)])
    grad = torch.zeros_like(Gn)
    inner = tuple(slice(1, n - 1) for n in grid.shape)
    grad[inner] = -Gn[inner] / kc.double()[inner]
    return grad


SEG_FASTEST = -1  # WFD_SEG_FASTEST: keep every level when it fits (see GradientPlan)


class GradientPlan:
    """Checkpointed forward + adjoint for ``nsteps`` steps from rest (``wfd_plan_*``).

    ``seg`` is the checkpoint spacing (0 = memory-optimal, about sqrt(2 nsteps); ``SEG_FASTEST``
    = every level kept, no recompute, when that store fits ``max_bytes`` or, with ``max_bytes``
    0, 0.8 of the free device memory, else memory-optimal); ``max_bytes`` caps the device memory
    of the plan.  The gradient does not depend on the spacing.
    """

    def __init__(self, grid: WaveGrid, nsteps: int, seg: int = 0, max_bytes: int = 0):
        self.grid, self.nsteps = grid, int(nsteps)
        lib = _lib.load()
        self._p = ctypes.c_void_p()
        check(lib.wfd_plan_create(ctypes.byref(self._p), ctypes.byref(grid.c), self.nsteps, int(seg),
                                  int(max_bytes)))
        s, k, b = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_size_t()
        check(lib.wfd_plan_info(self._p, ctypes.byref(s), ctypes.byref(k), ctypes.byref(b)))
        self.seg, self.nckpt, self.bytes = s.value, k.value, b.value

    def __del__(self):
        p = getattr(self, "_p", None)
        if p is not None and p.value:
            try:
                _lib.load().wfd_plan_destroy(p)
            except Exception:
                pass
            self._p = None

    def forward(self, medium: Medium, src: Optional[PointSet], sig, rec: Optional[PointSet],
                keep_state=False):
        """Traces ``(nshots, nsteps, R)``; checkpoints kept for :meth:`backward`."""
        g = self.grid
        s_sig = _sig_tensor(g, src, sig, self.nsteps)
        traces = torch.zeros(g.nshots, self.nsteps, rec.count if rec else 0, dtype=g.dtype, device=g.device)
        sp = src.struct(s_sig) if src is not None else None
        rp = rec.struct(traces) if rec is not None and rec.count else None
        last = g.zeros(nshots=2 * g.nshots) if keep_state else None
        check(_lib.load().wfd_plan_forward(self._p, ctypes.byref(medium.c), ctypes.byref(sp) if sp else None,
                                           ctypes.byref(rp) if rp else None, _ptr(last),
                                           _stream_ptr(g.device)))
        self._sig = s_sig
        if keep_state:
            return traces, (last[: g.nshots], last[g.nshots:])
        return traces

    def backward(self, medium: Medium, src: Optional[PointSet], rec: PointSet, resid) -> torch.Tensor:
        """``dJ/dkc`` on the nodes (float64, summed over shots) for ``dJ/dd = resid``.

        The adjoint injects ``-resid`` at the receivers; ``G = sum_n mu^{n+1} (L u^n + s^n)``
        and ``dJ/dkc = -G / kc``.
        """
        g = self.grid
        adj_sig = (-torch.as_tensor(resid, device=g.device)).to(g.dtype).contiguous()
        if adj_sig.shape != (g.nshots, self.nsteps, rec.count):
            raise ValueError(f"residual must be (nshots, nsteps, receivers) = {(g.nshots, self.nsteps, rec.count)}")
        ap = rec.struct(adj_sig)
        sp = src.struct(self._sig) if src is not None and src.count else None
        mu_src = None
        if src is not None and src.count:
            mu_src = torch.zeros(g.nshots, self.nsteps, src.count, dtype=g.dtype, device=g.device)
        mp = src.struct(mu_src) if mu_src is not None else None
        G = torch.zeros(g.elems, dtype=torch.float64, device=g.device)
        check(_lib.load().wfd_plan_backward(self._p, ctypes.byref(medium.c), ctypes.byref(sp) if sp else None,
                                            ctypes.byref(ap) if ap else None, ctypes.byref(mp) if mp else None,
                                            _ptr(G), _stream_ptr(g.device)))
        Gn = g.nodes(G[None])[0].clone()
        self.last_G = Gn.clone()  # sum_n mu^{n+1} L u^n before the source term
        self.last_mu_src = mu_src
        kc = medium.kc_nodes()
        if kc.shape[0] != 1:
            raise ValueError("the summed gradient needs one medium shared by all shots")
        return kc_gradient(g, Gn, mu_src, self._sig, src, kc[0])
