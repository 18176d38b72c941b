"""Continuous-batching solve server on one GPU (no counterpart in the reference).

Requests (heat or Jacobi solves of mixed 2D/3D sizes, see :class:`pyheat3d_amd.batch.Request`)
arrive over time. The serving loop, on one host thread driving one HIP stream:

1. admits arrived requests in arrival order while the device bytes of all resident solves stay
   within an HBM budget (fields are generated on the device from the request seed and stay
   resident until the solve finishes);
2. when the resident set changed, uploads the solve table (``ph3_batch_set``) and re-records the
   round graph;
3. replays one round: every resident solve advances by up to ``round_sweeps`` sweeps, small
   grids in the persistent workgroup-per-grid kernel and large grids in varlen tile passes, so
   new requests start while older ones are still running and their work shares the device;
4. retires finished solves (latency = completion - arrival), frees their memory, repeats.

A solve's result depends only on the solve, never on what else was resident, so it is the same
alone and in the mix (``tests/test_server_gpu.py`` checks this bit for bit).
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _hip
from .batch import (STATUS, Batch, DeviceFields, Request, Result, apply_tuned_tiles,
                    current_buffer, make_desc)


def poisson_arrivals(n: int, rate: float, seed: int = 7) -> np.ndarray:
    """Arrival times (s) of a Poisson process of ``rate`` requests per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


@dataclass
class ServeStats:
    solves: int
    seconds: float
    solves_per_s: float
    p50_ms: float
    p99_ms: float
    max_ms: float
    rounds: int
    max_resident: int
    max_resident_bytes: int
    rejected: int
    latencies_ms: list[float] = field(default_factory=list, repr=False)
    table_updates: int = 0        # uploads of the solve table (admissions / retirements)
    table_update_s: float = 0.0   # host time of those uploads and the round-graph re-records


class Server:
    """Admission by HBM budget, resident fields, round-based continuous batching."""

    def __init__(self, hbm_budget_bytes: int = 200 * 10 ** 9, round_sweeps: int = 128,
                 capacity: int = 4096, use_graph: bool = True,
                 keep_fields: bool | set[int] = False):
        """``keep_fields``: download every final field, or those of the request ids in a set."""
        self.budget = hbm_budget_bytes
        self.round_sweeps = round_sweeps
        self.use_graph = use_graph
        self.keep_fields = keep_fields
        self.stream = _hip.Stream()
        self.batch = Batch(capacity, self.stream)
        from .batch import default_params

        for pr in default_params():
            self.batch.params(*pr)
        self.capacity = capacity

    def close(self) -> None:
        self.batch.close()
        self.stream.close()

    def serve(self, requests: list[Request], arrivals: np.ndarray, time_limit: float = 600.0,
              progress=None, every_s: float = 30.0) -> tuple[list[Result], ServeStats]:
        """Serve ``requests`` arriving at ``arrivals`` (seconds after start); blocking.
        ``progress(rounds, finished, resident, seconds)`` is called every ``every_s`` seconds."""
        apply_tuned_tiles(requests)
        order = np.argsort(arrivals, kind="stable")
        pending = [(float(arrivals[i]), requests[i]) for i in order]
        for _, r in pending:
            if r.device_bytes() > self.budget:
                raise ValueError(f"request {r.rid} needs {r.device_bytes()} B > HBM budget")
        resident: list[tuple[Request, DeviceFields, float]] = []
        descs: list[_hip.SolveDesc] = []
        results: dict[int, Result] = {}
        used = 0
        graph = None
        dirty = True
        rounds = max_res = max_bytes = updates = 0
        t_update = 0.0
        t0 = time.perf_counter()
        t_last = t0
        try:
            while pending or resident:
                now = time.perf_counter() - t0
                if now > time_limit:
                    raise TimeoutError(f"serving exceeded {time_limit} s")
                # 1. admission
                while (pending and pending[0][0] <= now and len(resident) < self.capacity
                       and used + pending[0][1].device_bytes() <= self.budget):
                    arr, req = pending.pop(0)
                    f = DeviceFields(req, self.stream)
                    resident.append((req, f, arr))
                    descs.append(make_desc(req, f))
                    used += req.device_bytes()
                    dirty = True
                if not resident:
                    # idle until the next arrival
                    time.sleep(max(0.0, min(pending[0][0] - now, 0.001)))
                    continue
                max_res = max(max_res, len(resident))
                max_bytes = max(max_bytes, used)
                # 2. table upload and round graph
                if dirty:
                    tu = time.perf_counter()
                    self.stream.sync()
                    # the round graph was recorded against the previous table: destroy it
                    # before the table changes under it
                    if graph is not None:
                        graph.close()
                        graph = None
                    self.batch.set(descs)
                    if self.use_graph:
                        graph = _hip.Graph(self.stream, lambda: self.batch.round(self.round_sweeps))
                    dirty = False
                    updates += 1
                    t_update += time.perf_counter() - tu
                # 3. one round
                if graph is not None:
                    graph.launch(self.stream)
                else:
                    self.batch.round(self.round_sweeps)
                rounds += 1
                descs = self.batch.get()
                t_done = time.perf_counter() - t0
                # 4. retire finished solves
                keep_r, keep_d = [], []
                for (req, f, arr), d in zip(resident, descs):
                    if d.status == 0:
                        keep_r.append((req, f, arr))
                        keep_d.append(d)
                        continue
                    rel = None
                    if req.mode == "jacobi":
                        rel = math.sqrt(d.res_sq / d.res0_sq) if d.res0_sq > 0 else 0.0
                    res = Result(req.rid, STATUS[d.status], int(d.done), rel,
                                 seconds=t_done - arr)
                    if self.keep_fields is True or (self.keep_fields and
                                                    req.rid in self.keep_fields):
                        res.u = f.download(current_buffer(d), self.stream)
                    f.close()
                    used -= req.device_bytes()
                    results[req.rid] = res
                    dirty = True
                resident, descs = keep_r, keep_d
                if progress is not None and time.perf_counter() - t_last >= every_s:
                    t_last = time.perf_counter()
                    progress(rounds, len(results), len(resident), t_last - t0)
        finally:
            if graph is not None:
                graph.close()
            for _, f, _ in resident:
                f.close()
        total = time.perf_counter() - t0
        lat = np.array([results[r.rid].seconds for r in requests]) * 1e3
        stats = ServeStats(len(requests), total, len(requests) / total,
                           float(np.percentile(lat, 50)), float(np.percentile(lat, 99)),
                           float(lat.max()), rounds, max_res, max_bytes, 0, lat.tolist(),
                           updates, t_update)
        return [results[r.rid] for r in requests], stats
