"""Batched solves of many uneven 2D/3D grids on one GPU, and the one-launch-per-solve baseline.

No counterpart in the reference (it advances one grid per ``Stepper``; the nearest thing is the
size loop of ``benchmark.run``, ``src/pyheat3d/benchmark.py:276-304``). A :class:`Request` is one
solve: FTCS heat for a fixed number of steps, or Jacobi (Laplace or Poisson) until the FP64
residual falls below ``tol`` relative to the initial one, with stagnation and a sweep cap.

:class:`BatchSolver` keeps every field resident in HBM and advances all of them with the
library's batched rounds (``ph3_batch_round``): a persistent workgroup-per-grid kernel for small
grids and varlen tile passes for large ones, recorded once into a HIP graph and replayed until
every solve is done. :func:`solve_one_by_one` is the comparator: the same requests, each run on
its own ``HipStepper`` with one kernel launch per sweep, one after another.

Every solve's result depends only on the solve (tier, tile shapes and residual check points are
functions of its shape and mode), so a solve gives the same bits alone or in any batch.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _hip
from .jacobi import jacobi_rate
from .problem import heat_stencil, jacobi_stencil

STATUS = {0: "running", 1: "done", 2: "stagnated", 3: "max_sweeps"}


@dataclass
class Request:
    """One solve. Fields are generated on the device from ``seed`` (``ph3_fill_random``)."""

    shape: tuple[int, ...]
    mode: str = "heat"            # "heat" | "jacobi"
    storage: str = "float32"
    steps: int = 200              # heat: steps; jacobi: sweep cap
    r: float | None = None        # heat: mesh Fourier number (default 0.2 in 2D, 0.15 in 3D)
    tol: float = 1e-6             # jacobi: relative residual target
    seed: int = 0
    stall_window: int = 20
    stall_fraction: float = 0.05
    source: np.ndarray | None = None   # optional float32 source (Poisson)
    rid: int = 0
    bc: str = "dirichlet"         # "dirichlet" | "periodic" (heat only)
    tile: int = 0                 # tier-2 tile shape of fused passes (0: the library's rule)

    def __post_init__(self) -> None:
        if self.bc not in ("dirichlet", "periodic"):
            raise ValueError(f"bc must be 'dirichlet' or 'periodic', got {self.bc!r}")
        if self.bc == "periodic" and self.mode == "jacobi":
            raise ValueError("periodic Jacobi has no unique solution; use bc='dirichlet'")

    @property
    def dims(self) -> int:
        return len(self.shape)

    @property
    def sites(self) -> int:
        return int(np.prod(self.shape))

    @property
    def kind(self) -> int:
        """Initial state kind of ``ph3_fill_random``: heat 0 (random inside), Laplace 1."""
        return 0 if self.mode == "heat" else 1

    def stencil(self):
        if self.mode == "heat":
            r = self.r if self.r is not None else (0.2 if self.dims == 2 else 0.15)
            return heat_stencil(self.dims, r, self.bc)
        return jacobi_stencil(self.dims)

    def device_bytes(self) -> int:
        b = 2 * self.sites * _hip.STORAGE_BYTES[self.storage]
        return b + (self.sites * 4 if self.source is not None else 0)

    def work(self) -> int:
        """Lattice updates if the solve runs to its cap (a bound for Jacobi)."""
        inner = 0 if self.bc == "periodic" else 2
        return int(np.prod([s - inner for s in self.shape])) * self.steps


@dataclass
class Result:
    rid: int
    status: str
    steps: int
    rel_residual: float | None
    seconds: float | None = None
    u: np.ndarray | None = None
    extra: dict = field(default_factory=dict)


def make_mix(n: int = 2048, seed: int = 2024, frac2d: float = 0.6,
             n2: tuple[int, int] = (64, 2048), n3: tuple[int, int] = (32, 256),
             frac_jacobi: float = 0.5, heat_steps: int = 200, jacobi_cap: int = 10 ** 6,
             storage: str = "float32") -> list[Request]:
    """The mixed request set: 2D/3D, sizes log-uniform, heat or Jacobi, one seed per request."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        dims = 2 if rng.random() < frac2d else 3
        lo, hi = n2 if dims == 2 else n3
        N = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        jac = rng.random() < frac_jacobi
        out.append(Request(shape=(N,) * dims, mode="jacobi" if jac else "heat",
                           storage=storage, steps=jacobi_cap if jac else heat_steps,
                           seed=int(rng.integers(1, 2 ** 62)), rid=i))
    return out


class DeviceFields:
    """Two resident storage buffers (and an optional float32 source) for one request."""

    def __init__(self, req: Request, stream: _hip.Stream):
        self.req = req
        nb = req.sites * _hip.STORAGE_BYTES[req.storage]
        self.buf = [_hip.malloc(nb), _hip.malloc(nb)]
        self.src = 0
        code = _hip.STORAGE_CODE[req.storage]
        nz, ny, nx = (1, *req.shape) if req.dims == 2 else req.shape
        for p in self.buf:
            _hip.check(_hip.lib().ph3_fill_random(code, p, req.dims, nz, ny, nx, req.seed,
                                                  req.kind, stream.handle), "fill_random")
        if req.source is not None:
            s = np.ascontiguousarray(req.source, dtype=np.float32)
            self.src = _hip.malloc(s.nbytes)
            _hip.memcpy(self.src, s.ctypes.data, s.nbytes, _hip.H2D)

    def download(self, which: int, stream: _hip.Stream) -> np.ndarray:
        req = self.req
        out = np.empty(req.shape, np.float32)
        stream.sync()
        if req.storage == "float32":
            _hip.memcpy(out.ctypes.data, self.buf[which], out.nbytes, _hip.D2H)
            return out
        tmp = _hip.malloc(out.nbytes)
        try:
            _hip.check(_hip.lib().ph3_decode(_hip.STORAGE_CODE[req.storage], self.buf[which], tmp,
                                             req.sites, stream.handle), "decode")
            stream.sync()
            _hip.memcpy(out.ctypes.data, tmp, out.nbytes, _hip.D2H)
        finally:
            _hip.free(tmp)
        return out

    def close(self) -> None:
        for p in self.buf + [self.src]:
            _hip.free(p)
        self.buf, self.src = [0, 0], 0


def apply_tuned_tiles(requests: list[Request]) -> None:
    """Give every request without a tile shape the tuned one for its dims and size."""
    from .tune import batch_tile

    for r in requests:
        if r.tile == 0:
            r.tile = batch_tile(r.dims, r.shape)


def default_params() -> list[tuple[int, int, int, int]]:
    """The tuned chunk policies of the batch classes (none when the table holds none)."""
    from .tune import batch_policy

    return batch_policy()


def make_desc(req: Request, f: DeviceFields) -> _hip.SolveDesc:
    st = req.stencil()
    d = _hip.SolveDesc()
    d.dims = req.dims
    d.nz, d.ny, d.nx = (1, *req.shape) if req.dims == 2 else req.shape
    d.periodic = 1 if req.bc == "periodic" else 0
    d.storage = _hip.STORAGE_CODE[req.storage]
    d.c0, d.c1 = float(np.float32(st.c0)), float(np.float32(st.c1))
    d.buf[0], d.buf[1] = f.buf
    d.src = f.src or None
    d.mode = 0 if req.mode == "heat" else 1
    d.tile = req.tile
    d.steps = req.steps
    d.tol = req.tol
    d.stall_window = req.stall_window if req.mode == "jacobi" else 0
    d.stall_fraction = req.stall_fraction
    d.log_rate = math.log(max(jacobi_rate(req.shape), 1e-300)) * _hip.BATCH_CHECK
    return d


class Batch:
    """A device solve table (``ph3_batch``) bound to one stream."""

    def __init__(self, capacity: int, stream: _hip.Stream):
        if _hip.lib().ph3_solve_desc_bytes() != ctypes.sizeof(_hip.SolveDesc):
            raise _hip.HipError("ph3_solve_desc layout mismatch between library and binding")
        p = ctypes.c_void_p()
        _hip.check(_hip.lib().ph3_batch_create(ctypes.byref(p), capacity), "batch_create")
        self.handle = int(p.value or 0)
        self.stream = stream
        self.n = 0

    def params(self, dims: int, kpass: int, target_tiles: int, max_chunk: int) -> None:
        """March-chunk policy of the (dims, kpass) classes from the next :meth:`set` on."""
        _hip.check(_hip.lib().ph3_batch_params(self.handle, dims, kpass, target_tiles,
                                               max_chunk), "batch_params")

    def set(self, descs: list[_hip.SolveDesc]) -> None:
        arr = (_hip.SolveDesc * max(1, len(descs)))(*descs)
        _hip.check(_hip.lib().ph3_batch_set(self.handle, arr, len(descs), self.stream.handle),
                   "batch_set")
        self.n = len(descs)

    def round(self, sweeps: int) -> None:
        _hip.check(_hip.lib().ph3_batch_round(self.handle, sweeps, self.stream.handle),
                   "batch_round")

    def get(self) -> list[_hip.SolveDesc]:
        arr = (_hip.SolveDesc * max(1, self.n))()
        _hip.check(_hip.lib().ph3_batch_get(self.handle, arr, self.n, self.stream.handle),
                   "batch_get")
        return list(arr)[: self.n]

    def close(self) -> None:
        if self.handle:
            _hip.check(_hip.lib().ph3_batch_destroy(self.handle), "batch_destroy")
            self.handle = 0


def current_buffer(d: _hip.SolveDesc) -> int:
    """Index of the buffer holding a solve's state: passes of ``kpass`` steps alternate them."""
    return (int(d.done) // max(1, int(d.kpass))) & 1


def _result(req: Request, d: _hip.SolveDesc) -> Result:
    rel = None
    if req.mode == "jacobi":
        rel = math.sqrt(d.res_sq / d.res0_sq) if d.res0_sq > 0 else 0.0
    return Result(req.rid, STATUS[d.status], int(d.done), rel)


class BatchSolver:
    """All requests resident at once, advanced by graph-replayed batched rounds."""

    def __init__(self, requests: list[Request], round_sweeps: int = 256, use_graph: bool = True,
                 compact: bool = True, params: list[tuple[int, int, int, int]] | None = None,
                 tuned_tiles: bool = True):
        """``params``: (dims, kpass, target_tiles, max_chunk) chunk policies, see
        :meth:`Batch.params` (default: the tuning table's, :func:`pyheat3d_amd.tune.batch_policy`).
        ``tuned_tiles``: requests without a tile shape get the table's (else the library's rule)."""
        self.requests = requests
        self.round_sweeps = round_sweeps
        self.use_graph = use_graph
        self.compact = compact
        self.stream = _hip.Stream()
        self.fields: list[DeviceFields] = []
        self.batch: Batch | None = None
        if tuned_tiles:
            apply_tuned_tiles(requests)
        try:
            for r in requests:
                self.fields.append(DeviceFields(r, self.stream))
            self.stream.sync()
            self.batch = Batch(len(requests), self.stream)
            for pr in (params if params is not None else default_params()):
                self.batch.params(*pr)
            self.state = [make_desc(r, f) for r, f in zip(requests, self.fields)]
            self.table = list(range(len(requests)))  # request index of each table row
            self.batch.set(self.state)
        except Exception:
            self.close()
            raise

    def run(self, max_rounds: int = 10 ** 6, progress=None, every_s: float = 30.0) -> dict:
        """Replay rounds until every solve has finished; returns timing and round counts.

        Every request counts as submitted when the run starts; its completion time is the end
        of the round in which it finished (``finish_s``, seconds from the start).

        When half of the table has finished, the finished solves are dropped from the device
        table (their state is kept on the host) and the round graph is re-recorded, so the
        passes of the remaining solves do not walk the tiles of finished ones.
        """
        graph = None
        t0 = time.perf_counter()
        t_last = t0
        rounds = compactions = 0
        self.finish_s: list[float | None] = [None] * len(self.requests)
        try:
            while rounds < max_rounds:
                if graph is None and self.use_graph:
                    graph = _hip.Graph(self.stream, lambda: self.batch.round(self.round_sweeps))
                if graph is not None:
                    graph.launch(self.stream)
                else:
                    self.batch.round(self.round_sweeps)
                rounds += 1
                descs = self.batch.get()
                t_round = time.perf_counter() - t0
                for row, d in zip(self.table, descs):
                    self.state[row] = d
                    if d.status != 0 and self.finish_s[row] is None:
                        self.finish_s[row] = t_round
                running = [row for row, d in zip(self.table, descs) if d.status == 0]
                if progress is not None and time.perf_counter() - t_last >= every_s:
                    t_last = time.perf_counter()
                    progress(rounds, len(running), t_last - t0)
                if not running:
                    break
                if self.compact and len(running) <= len(self.table) // 2:
                    # the round graph was recorded against the current table: destroy it
                    # before the table changes under it
                    if graph is not None:
                        graph.close()
                        graph = None
                    self.table = running
                    self.batch.set([self.state[row] for row in running])
                    compactions += 1
        finally:
            if graph is not None:
                graph.close()
        return {"seconds": time.perf_counter() - t0, "rounds": rounds, "compactions": compactions}

    def results(self, keep_fields: bool | set[int] = False) -> list[Result]:
        """One result per request; ``keep_fields`` downloads every final field, or the fields
        of the request ids in a set."""
        out = []
        finish = getattr(self, "finish_s", [None] * len(self.requests))
        for req, f, d, t in zip(self.requests, self.fields, self.state, finish):
            res = _result(req, d)
            res.seconds = t
            if keep_fields is True or (keep_fields and req.rid in keep_fields):
                res.u = f.download(current_buffer(d), self.stream)
            out.append(res)
        return out

    def close(self) -> None:
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        for f in self.fields:
            f.close()
        self.fields = []
        self.stream.close()


def solve_one_by_one(requests: list[Request], keep_fields: bool = False,
                     check_every: int = _hip.BATCH_CHECK, graphs: bool = False,
                     kpass: dict[int, int] | None = None,
                     progress=None) -> tuple[list[Result], float]:
    """Baseline: each request on its own stepper, one after another; one launch per pass (the
    reference's stepping), or with ``graphs`` the stepper's replayed graphs of 16 passes for
    grids of up to ``HipStepper.GRAPH_MAX_SITES`` sites.

    ``kpass`` (request id -> steps fused per pass, as the batch chose them: the ``kpass`` field
    of its solve table) makes each solve run the same rung as in the batch, the k-step
    temporal-blocking kernel (same float32 bits as k single steps), so that a batched-vs-per-
    solve speedup measures batching, not temporal blocking. Without it every solve runs the
    single-step march rung (z-march in 3D, y-march in 2D).

    Jacobi requests use the same residual definition, check cadence, tolerance, stagnation test
    and cap as the batch (:func:`pyheat3d_amd.jacobi.solve`).
    """
    from .backends.hip_backend import make
    from .jacobi import solve

    t_all = time.perf_counter()
    out = []
    for req in requests:
        k = (kpass or {}).get(req.rid, 1)
        if k > 1:
            cls = make("tb", req.storage, k)
        else:   # the single-grid march rungs: z-march in 3D, y-march in 2D
            cls = make("zmarch", req.storage) if req.dims == 3 else make("tb", req.storage, 1)
        stream = _hip.Stream()
        f = DeviceFields(req, stream)
        stream.sync()
        u0 = f.download(0, stream)
        f.close()
        stream.close()
        t0 = time.perf_counter()
        st = cls(u0, req.stencil(), req.source)
        if not graphs:
            st.GRAPH_MAX_SITES = 0
        try:
            if req.mode == "heat":
                st.step(req.steps)
                st.synchronize()
                res = Result(req.rid, "done", req.steps, None)
            else:
                jr = solve("", u0, req.source, tol=req.tol, check_every=check_every,
                           max_sweeps=req.steps, stall_window=req.stall_window,
                           stall_fraction=req.stall_fraction, keep=False, stepper=st)
                res = Result(req.rid, {"converged": "done"}.get(jr.status, jr.status),
                             jr.sweeps, jr.rel_residual)
            if keep_fields:
                res.u = st.get()
        finally:
            st.close()
        res.seconds = time.perf_counter() - t0
        out.append(res)
        if progress is not None:
            progress(len(out), len(requests), time.perf_counter() - t_all)
    return out, time.perf_counter() - t_all


def batch_kpass(solver: BatchSolver) -> dict[int, int]:
    """Steps fused per pass that the batch chose for each request (its solve table)."""
    return {req.rid: max(1, int(d.kpass)) for req, d in zip(solver.requests, solver.state)}


def error_vs_fp64(req: Request, u: np.ndarray) -> float:
    """Max-norm relative error of a final field against the FP64 NumPy solution of the same
    request (:mod:`pyheat3d_amd.fp64`): heat, the same steps in float64; Jacobi, the fixed point
    of the device's Jacobi map solved exactly in float64."""
    from . import fp64
    from .analytic import splitmix_field

    u0 = splitmix_field(req.shape, req.seed, req.kind)
    st = req.stencil()
    if req.mode == "heat":
        ref = fp64.heat(u0, st.c0, st.c1, req.bc, req.steps)
    else:
        ref = fp64.dirichlet_solution(u0, st.c1, req.source)
    return fp64.rel_err(u, ref)

