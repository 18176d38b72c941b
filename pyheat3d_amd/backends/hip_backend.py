"""``hip-cdna4/*`` variants: hand-written HIP kernels for gfx950 behind the Stepper interface.

Counterpart of the reference's ``CupyStepper`` (``src/pyheat3d/backends/cupy_backend.py:115-213``):
same kinds (naive / zmarch / tb), same pass schedule (``n // k`` fused passes, then ``n % k``
naive single steps, ping-pong after each, ``cupy_backend.py:183-194``), same storage handling
(float32 arithmetic, values rounded to the storage type once per pass). Differences: kernels are
precompiled into ``libpyheat3d_hip.so`` and called through its C ABI; the storage codec runs on
the device; the source term ``b`` is supported by every kernel; each stepper owns a HIP stream.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import _hip
from ..problem import Stencil
from .base import Stepper


class HipStepper(Stepper):
    """``kind`` in {"naive", "zmarch", "tb"}; ``K`` = steps fused per pass for "tb"."""

    kind = "naive"
    K = 1
    storage = "float32"
    #: optional ``_hip.Tile`` override of the launch shape (the autotuner sets it)
    tile_override: _hip.Tile | None = None

    def __init__(self, u0: np.ndarray, stencil: Stencil, source: np.ndarray | None = None):
        super().__init__(u0, stencil, source)
        if _hip.device_count() < 1:
            raise RuntimeError("no HIP device")
        self.steps_per_pass = self.K if self.kind == "tb" else 1
        self._kind_code = {"naive": _hip.NAIVE, "zmarch": _hip.ZMARCH, "tb": _hip.TB}[self.kind]
        if self.kind == "zmarch" and self.dims != 3 and self.K == 1 and not self._allow_2d:
            raise ValueError("zmarch is 3D only (use ymarch/tb in 2D)")
        self._c0 = float(np.float32(self.c0))
        self._c1 = float(np.float32(self.c1))
        self.n = int(np.prod(self.shape))
        self.nbytes = self.n * _hip.STORAGE_BYTES[self.storage]
        self.stream = _hip.Stream()
        self.a = self.b = self.src = 0
        self._graphs: dict = {}
        try:
            self.a = _hip.malloc(self.nbytes)
            self.b = _hip.malloc(self.nbytes)
            self._upload(np.ascontiguousarray(u0, dtype=np.float32))
            if source is not None:
                s = np.ascontiguousarray(source, dtype=np.float32)
                self.src = _hip.malloc(s.nbytes)
                _hip.memcpy(self.src, s.ctypes.data, s.nbytes, _hip.H2D)
        except Exception:
            self.close()
            raise
        if self.tile_override is None and self.kind != "naive":
            from ..tune import lookup

            self.tile_override = lookup(self.kind, self.steps_per_pass, self.dims, self.storage,
                                        self.bc, self.shape, source is not None)
        self.tile = self.plan()

    def plan(self) -> dict:
        """The launch of this stepper's main pass as the library resolves it (lane width, rows,
        waves, march chunk, tile geometry, workgroups), and the tuned tile it started from."""
        p = _hip.sweep_plan(self._kind_code, self.steps_per_pass, self.storage, self.periodic,
                            self.shape, self.a, self.b, self.src or None, self.tile_override)
        return {**p, "tuned": self.tile_override.as_dict() if self.tile_override else None}

    _allow_2d = False

    # -- host <-> device ---------------------------------------------------------------------
    def _upload(self, host: np.ndarray) -> None:
        # Every copy is ordered on this stepper's stream and completed before the first sweep:
        # the stream is non-blocking, so it does not wait for work on the null stream.
        code = _hip.STORAGE_CODE[self.storage]
        s = self.stream
        if self.storage == "float32":
            _hip.memcpy_async(self.a, host.ctypes.data, host.nbytes, _hip.H2D, s)
        else:
            tmp = _hip.malloc(host.nbytes)
            try:
                _hip.memcpy_async(tmp, host.ctypes.data, host.nbytes, _hip.H2D, s)
                _hip.check(_hip.lib().ph3_encode(code, tmp, self.a, self.n, s.handle), "encode")
                s.sync()
            finally:
                _hip.free(tmp)
        # both ping-pong buffers carry the (fixed) Dirichlet boundary
        _hip.memcpy_async(self.b, self.a, self.nbytes, _hip.D2D, s)
        s.sync()

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float32)
        self.stream.sync()
        if self.storage == "float32":
            _hip.memcpy(out.ctypes.data, self.a, out.nbytes, _hip.D2H)
            return out
        tmp = _hip.malloc(out.nbytes)
        try:
            _hip.check(_hip.lib().ph3_decode(_hip.STORAGE_CODE[self.storage], self.a, tmp, self.n,
                                             self.stream.handle), "decode")
            self.stream.sync()
            _hip.memcpy(out.ctypes.data, tmp, out.nbytes, _hip.D2H)
        finally:
            _hip.free(tmp)
        return out

    # -- stepping ----------------------------------------------------------------------------
    def _launch(self, kind: int, k: int) -> None:
        # the tile override tunes the main kernel; remainder steps use the naive defaults
        tile = self.tile_override if kind == self._kind_code else None
        _hip.sweep(kind, k, self.storage, self.periodic, self.shape, self._c0, self._c1,
                   self.a, self.b, self.src or None, tile, self.stream.handle)
        self.a, self.b = self.b, self.a

    #: grids of up to this many sites replay captured passes: their sweeps take a few
    #: microseconds, comparable to launching one from Python (MI355X: 64^3 naive 56 -> 83,
    #: 1024^2 y-march 213 -> 317 GLUPS; from 2048^2 on graphs measured 1-3% slower)
    GRAPH_MAX_SITES = 1 << 20
    #: passes per captured graph (even, so the ping-pong buffers end where they began)
    GRAPH_PASSES = 16

    def step(self, n: int = 1) -> None:
        k = self.steps_per_pass
        passes = n // k
        if self.n <= self.GRAPH_MAX_SITES and passes > self.GRAPH_PASSES:
            self._launch(self._kind_code, k)   # direct first: library caches fill outside capture
            passes -= 1
            g = self._pass_graph()
            for _ in range(passes // self.GRAPH_PASSES):
                g.launch(self.stream)
            passes %= self.GRAPH_PASSES
        for _ in range(passes):
            self._launch(self._kind_code, k)
        for _ in range(n % k):
            self._launch(_hip.NAIVE, 1)
        self.steps_done += n

    def _pass_graph(self) -> _hip.Graph:
        """HIP graph of GRAPH_PASSES main passes from the current buffer order and tile (cached:
        the same kernels with the same arguments as direct launches)."""
        tile = self.tile_override.as_dict() if self.tile_override else None
        key = (self.a, self.b, tuple(sorted(tile.items())) if tile else None)
        g = self._graphs.get(key)
        if g is None:
            def record():
                for _ in range(self.GRAPH_PASSES):
                    self._launch(self._kind_code, self.steps_per_pass)
            a, b = self.a, self.b
            try:
                g = self._graphs[key] = _hip.Graph(self.stream, record)
            finally:   # recording swaps the buffers an even number of times, or fails midway
                self.a, self.b = a, b
        return g

    def synchronize(self) -> None:
        self.stream.sync()

    _work = 0

    def residual_sq(self) -> float:
        """FP64 squared 2-norm of ``b + c0 u + c1 S(u) - u`` for the current state (blocking)."""
        if not self._work:  # result at offset 0, workgroup partials from offset 256
            self._work = _hip.malloc(int(_hip.lib().ph3_residual_work_bytes()) + 256)
        _hip.check(_hip.lib().ph3_residual(
            _hip.STORAGE_CODE[self.storage], int(self.periodic), self.dims,
            *((1, *self.shape) if self.dims == 2 else self.shape), self._c0, self._c1,
            self.a, self.src or None, self._work, self._work + 256, self.stream.handle),
            "residual")
        self.stream.sync()
        res = ctypes.c_double()
        _hip.memcpy(ctypes.addressof(res), self._work, 8, _hip.D2H)
        return float(res.value)

    def close(self) -> None:
        for g in getattr(self, "_graphs", {}).values():
            g.close()
        self._graphs = {}
        for name in ("a", "b", "src", "_work"):
            p = getattr(self, name, 0)
            if p:
                _hip.free(p)
                setattr(self, name, 0)
        st = getattr(self, "stream", None)
        if st is not None:
            st.close()


def make(kind: str, storage: str = "float32", K: int = 1) -> type[HipStepper]:
    """HipStepper subclass for a kernel kind, storage type and K (like ``cupy_backend.make``)."""
    _hip.lib()  # raise now if the library is missing
    if kind == "tb" and K > 1 and not any(_hip.lib().ph3_tb_supported(d, K) for d in (2, 3)):
        raise RuntimeError(f"temporal blocking with k={K} is not built")
    attrs = {"kind": kind, "storage": storage, "K": K}
    if kind == "tb" and K == 1:  # ymarch: the K = 1 case runs the march kernel in 2D
        attrs.update(kind="zmarch", _allow_2d=True)
    return type(f"Hip_{kind}{K}_{storage}", (HipStepper,), attrs)
