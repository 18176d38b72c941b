"""ctypes binding of ``libpyheat3d_hip.so`` (declared in ``include/pyheat3d_hip.h``).

There is no fallback: if the library is missing or no gfx950 device is present, the calls
raise. Build the library with ``python -m pyheat3d_amd.build``.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_void_p
from pathlib import Path

LIB_PATH = Path(os.environ.get("PH3_LIB", Path(__file__).resolve().parent / "libpyheat3d_hip.so"))

NAIVE, ZMARCH, TB = 0, 1, 2
STORAGE_CODE = {"float32": 0, "float16": 1, "bfloat16": 2}
STORAGE_BYTES = {"float32": 4, "float16": 2, "bfloat16": 2}
H2D, D2H, D2D = 0, 1, 2


class HipError(RuntimeError):
    """A call into libpyheat3d_hip.so failed."""


class Tile(ctypes.Structure):
    """``ph3_tile``: rows per wave, march chunk, sites per lane, waves, flags (0 = default)."""

    _fields_ = [("ty", c_int), ("zch", c_int), ("vec", c_int), ("waves", c_int),
                ("flags", c_int)]

    def as_dict(self) -> dict[str, int]:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class Plan(ctypes.Structure):
    """``ph3_plan``: the launch a sweep resolves (lane width, rows, waves, chunk, tile, groups)."""

    _fields_ = [("kind", c_int), ("k", c_int), ("vec", c_int), ("rows", c_int),
                ("waves", c_int), ("zch", c_int), ("tile_x", c_int), ("tile_y", c_int),
                ("out_x", c_int), ("out_y", c_int), ("threads", c_int),
                ("groups", ctypes.c_longlong)]

    def as_dict(self) -> dict[str, int]:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


# name -> (restype, argtypes); every symbol of include/pyheat3d_hip.h
SIGNATURES: dict[str, tuple[object, list[object]]] = {
    "ph3_last_error": (c_char_p, []),
    "ph3_version": (c_int, []),
    "ph3_device_count": (c_int, [POINTER(c_int)]),
    "ph3_device_info": (c_int, [c_int, c_char_p, c_int, c_char_p, c_int, POINTER(c_int),
                                POINTER(c_size_t), POINTER(c_int), POINTER(c_int),
                                POINTER(c_int), POINTER(c_int)]),
    "ph3_set_device": (c_int, [c_int]),
    "ph3_device_sync": (c_int, []),
    "ph3_mem_info": (c_int, [POINTER(c_size_t), POINTER(c_size_t)]),
    "ph3_malloc": (c_int, [POINTER(c_void_p), c_size_t]),
    "ph3_free": (c_int, [c_void_p]),
    "ph3_host_alloc": (c_int, [POINTER(c_void_p), c_size_t]),
    "ph3_host_free": (c_int, [c_void_p]),
    "ph3_memcpy": (c_int, [c_void_p, c_void_p, c_size_t, c_int]),
    "ph3_memcpy_async": (c_int, [c_void_p, c_void_p, c_size_t, c_int, c_void_p]),
    "ph3_stream_create": (c_int, [POINTER(c_void_p)]),
    "ph3_stream_destroy": (c_int, [c_void_p]),
    "ph3_stream_sync": (c_int, [c_void_p]),
    "ph3_event_create": (c_int, [POINTER(c_void_p)]),
    "ph3_event_destroy": (c_int, [c_void_p]),
    "ph3_event_record": (c_int, [c_void_p, c_void_p]),
    "ph3_event_sync": (c_int, [c_void_p]),
    "ph3_event_elapsed_ms": (c_int, [POINTER(c_float), c_void_p, c_void_p]),
    "ph3_sweep": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_float,
                          c_float, c_void_p, c_void_p, c_void_p, POINTER(Tile), c_void_p]),
    "ph3_sweep_plan": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                               c_void_p, c_void_p, c_void_p, POINTER(Tile), POINTER(Plan)]),
    "ph3_tb_supported": (c_int, [c_int, c_int]),
    "ph3_encode": (c_int, [c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "ph3_decode": (c_int, [c_int, c_void_p, c_void_p, c_size_t, c_void_p]),
    "ph3_copy_kernel": (c_int, [c_void_p, c_void_p, c_size_t, c_void_p]),
    "ph3_fill_random": (c_int, [c_int, c_void_p, c_int, c_int, c_int, c_int, ctypes.c_ulonglong,
                                c_int, c_void_p]),
    "ph3_residual_work_bytes": (c_size_t, []),
    "ph3_residual": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_float, c_float,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph3_residual_noalloc_scratch_bytes": (c_size_t, []),
    "ph3_residual_noalloc": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_float, c_float,
                                     c_void_p, c_void_p, c_void_p, c_void_p]),
    "ph3_graph_begin": (c_int, [c_void_p]),
    "ph3_graph_end": (c_int, [c_void_p, POINTER(c_void_p)]),
    "ph3_graph_launch": (c_int, [c_void_p, c_void_p]),
    "ph3_graph_destroy": (c_int, [c_void_p]),
    "ph3_solve_desc_bytes": (c_size_t, []),
    "ph3_batch_create": (c_int, [POINTER(c_void_p), c_int]),
    "ph3_batch_destroy": (c_int, [c_void_p]),
    "ph3_batch_params": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "ph3_batch_set": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
    "ph3_batch_round": (c_int, [c_void_p, c_int, c_void_p]),
    "ph3_batch_get": (c_int, [c_void_p, c_void_p, c_int, c_void_p]),
}

BATCH_CHECK = 64          # PH3_BATCH_CHECK
BATCH_SMALL_SITES = 4096  # PH3_BATCH_SMALL_SITES
STALL_RING = 32           # PH3_STALL_RING


class SolveDesc(ctypes.Structure):
    """Mirror of ``ph3_solve_desc`` (size checked against the library at first use)."""

    _fields_ = [
        ("dims", c_int), ("nz", c_int), ("ny", c_int), ("nx", c_int),
        ("periodic", c_int), ("storage", c_int), ("c0", c_float), ("c1", c_float),
        ("buf", c_void_p * 2), ("src", c_void_p), ("mode", c_int), ("stall_window", c_int),
        ("kpass", c_int), ("tile", c_int), ("steps", ctypes.c_longlong), ("tol", c_double), ("log_rate", c_double),
        ("stall_fraction", c_double), ("done", ctypes.c_longlong), ("res0_sq", c_double),
        ("res_sq", c_double), ("status", c_int), ("checks", c_int),
        ("ring", c_double * STALL_RING),
    ]


class Graph:
    """An executable HIP graph captured from one stream."""

    def __init__(self, stream: "Stream", record) -> None:
        check(lib().ph3_graph_begin(stream.handle), "graph_begin")
        try:
            record()
        finally:
            p = c_void_p()
            rc = lib().ph3_graph_end(stream.handle, ctypes.byref(p))
        check(rc, "graph_end")
        self.handle = int(p.value or 0)

    def launch(self, stream: "Stream") -> None:
        check(lib().ph3_graph_launch(self.handle, stream.handle), "graph_launch")

    def close(self) -> None:
        if self.handle:
            check(lib().ph3_graph_destroy(self.handle), "graph_destroy")
            self.handle = 0

_lib: ctypes.CDLL | None = None


def lib() -> ctypes.CDLL:
    """Load the library once and declare every signature."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise HipError(f"{LIB_PATH} is not built (run: python -m pyheat3d_amd.build)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().ph3_last_error().decode(errors="replace")
        raise HipError(f"{what}: {msg} (code {rc})" if what else f"{msg} (code {rc})")


def device_count() -> int:
    n = c_int(0)
    check(lib().ph3_device_count(ctypes.byref(n)), "device_count")
    return n.value


def device_info(dev: int = 0) -> dict[str, object]:
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(256)
    cus, wave, lds, mclk, bus = c_int(), c_int(), c_int(), c_int(), c_int()
    mem = c_size_t()
    check(lib().ph3_device_info(dev, name, 256, arch, 256, ctypes.byref(cus), ctypes.byref(mem),
                                ctypes.byref(wave), ctypes.byref(lds), ctypes.byref(mclk),
                                ctypes.byref(bus)), "device_info")
    return {"name": name.value.decode(), "arch": arch.value.decode(), "cu_count": cus.value,
            "total_mem_bytes": mem.value, "wave_size": wave.value, "lds_per_cu": lds.value,
            "mem_clock_khz": mclk.value, "bus_width_bits": bus.value}


def malloc(nbytes: int) -> int:
    p = c_void_p()
    check(lib().ph3_malloc(ctypes.byref(p), nbytes), f"malloc({nbytes})")
    return int(p.value or 0)


def free(ptr: int) -> None:
    if ptr:
        check(lib().ph3_free(ptr), "free")


def memcpy(dst: int, src: int, nbytes: int, kind: int) -> None:
    """Blocking copy, complete on the device when it returns."""
    check(lib().ph3_memcpy(dst, src, nbytes, kind), "memcpy")


def memcpy_async(dst: int, src: int, nbytes: int, kind: int, stream: "Stream") -> None:
    """Copy ordered on ``stream`` (host buffers must stay alive until the stream is synced)."""
    check(lib().ph3_memcpy_async(dst, src, nbytes, kind, stream.handle), "memcpy_async")


def mem_info() -> tuple[int, int]:
    f, t = c_size_t(), c_size_t()
    check(lib().ph3_mem_info(ctypes.byref(f), ctypes.byref(t)), "mem_info")
    return f.value, t.value


class Stream:
    """An owned HIP stream (non-blocking, default priority)."""

    def __init__(self) -> None:
        p = c_void_p()
        check(lib().ph3_stream_create(ctypes.byref(p)), "stream_create")
        self.handle = int(p.value or 0)

    def sync(self) -> None:
        check(lib().ph3_stream_sync(self.handle), "stream_sync")

    def close(self) -> None:
        if self.handle:
            check(lib().ph3_stream_destroy(self.handle), "stream_destroy")
            self.handle = 0


class Event:
    def __init__(self) -> None:
        p = c_void_p()
        check(lib().ph3_event_create(ctypes.byref(p)), "event_create")
        self.handle = int(p.value or 0)

    def record(self, stream: int | None) -> None:
        check(lib().ph3_event_record(self.handle, stream or 0), "event_record")

    def sync(self) -> None:
        check(lib().ph3_event_sync(self.handle), "event_sync")

    def elapsed_ms(self, later: "Event") -> float:
        ms = c_float()
        check(lib().ph3_event_elapsed_ms(ctypes.byref(ms), self.handle, later.handle), "elapsed")
        return float(ms.value)

    def close(self) -> None:
        if self.handle:
            check(lib().ph3_event_destroy(self.handle), "event_destroy")
            self.handle = 0


def sweep(kind: int, k: int, storage: str, periodic: bool, shape: tuple[int, ...], c0: float,
          c1: float, u: int, o: int, b: int | None, tile: Tile | None, stream: int | None) -> None:
    dims = len(shape)
    nz, ny, nx = (1, *shape) if dims == 2 else shape
    rc = lib().ph3_sweep(kind, k, STORAGE_CODE[storage], int(periodic), dims, nz, ny, nx,
                         c0, c1, u, o, b or None, ctypes.byref(tile) if tile else None,
                         stream or None)
    check(rc, "sweep")


def sweep_plan(kind: int, k: int, storage: str, periodic: bool, shape: tuple[int, ...], u: int,
               o: int, b: int | None, tile: Tile | None) -> dict[str, int]:
    """The launch :func:`sweep` would make for these arguments (``ph3_sweep_plan``)."""
    dims = len(shape)
    nz, ny, nx = (1, *shape) if dims == 2 else shape
    plan = Plan()
    rc = lib().ph3_sweep_plan(kind, k, STORAGE_CODE[storage], int(periodic), dims, nz, ny, nx,
                              u, o, b or None, ctypes.byref(tile) if tile else None,
                              ctypes.byref(plan))
    check(rc, "sweep_plan")
    return plan.as_dict()

