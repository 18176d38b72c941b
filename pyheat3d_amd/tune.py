"""Autotuner for the ``hip-cdna4`` launch shapes, persisted as a tuning table.

Extends the reference's development sweep (``bench/tune.py:21-95``, which mutates module-level
tile dicts) into a table that the stepper reads: for each (arch, kind, k, dims, storage, bc,
size bucket) it stores the fastest ``ph3_tile`` (rows per wave, waves, lane vector, march
chunk, store flags). The size bucket is ``round(log2(sites))``. Every candidate is a shape the
library builds, and every shape computes the same bits (the per-site arithmetic does not depend
on the tile), so tuning can never change results.

    python -m pyheat3d_amd tune --dims 3 --n 1024 --kinds zmarch tb3
"""

from __future__ import annotations

import itertools
import json
import math
import os
from pathlib import Path
from typing import Any

import numpy as np

from . import _hip

TABLE = Path(os.environ.get("PH3_TUNING", Path(__file__).resolve().parent / "tuning" / "gfx950.json"))

ZMARCH_SHAPES = [(1, 16), (2, 8), (2, 4), (4, 4), (4, 8), (8, 2), (4, 2), (8, 1), (8, 8)]
TB3_SHAPES = [(2, 4, 16), (1, 8, 8), (2, 4, 8), (4, 2, 16), (1, 4, 16)]  # (vec, rows, waves)


def bucket(shape: tuple[int, ...]) -> int:
    return int(round(math.log2(max(1, int(np.prod(shape))))))


def table_key(kind: str, k: int, dims: int, storage: str, bc: str, shape: tuple[int, ...],
              source: bool = False) -> str:
    """Table key; a float32 source adds a register queue to tb, so sourced runs tune apart."""
    return f"{kind}/k{k}/{dims}d/{storage}/{bc}{'/src' if source else ''}/b{bucket(shape)}"


_table: dict[str, Any] | None = None


def load_table() -> dict[str, Any]:
    global _table
    if _table is None:
        _table = json.loads(TABLE.read_text()) if TABLE.exists() else {}
    return _table


def lookup(kind: str, k: int, dims: int, storage: str, bc: str,
           shape: tuple[int, ...], source: bool = False) -> _hip.Tile | None:
    """Tuned tile for this problem, or None (the library's defaults apply)."""
    e = load_table().get(table_key(kind, k, dims, storage, bc, shape, source))
    if not e:
        return None
    t = e["tile"]
    # an entry serves every shape of its size bucket; one whose lane vector does not divide this
    # row (tuned on 128^3, asked for 121^3) is not applied: the library's defaults pick a width
    if t["vec"] and shape[-1] % t["vec"]:
        return None
    return _hip.Tile(t["ty"], t["zch"], t["vec"], t["waves"], t.get("flags", 0))


def candidates(kind: str, k: int, dims: int, nx: int, storage: str = "float32") -> list[_hip.Tile]:
    # lane vectors up to 16 bytes: 4 float32 or 8 16-bit sites
    vecs = (1, 2, 4) if storage == "float32" else (1, 2, 4, 8)
    out = []
    if kind == "zmarch" and dims == 3:
        # chunks of 32-256 planes: sizes that are not powers of two need a chunk whose
        # workgroup count fills the last wave of workgroups
        for (r, w), v, z in itertools.product(ZMARCH_SHAPES, vecs, (0, 32, 48, 64, 96, 128, 256)):
            if nx % v == 0:
                out.append(_hip.Tile(r, z, v, w, 0))
    elif kind == "zmarch":
        for r, v, z in itertools.product((1, 2, 4, 8), vecs, (0, 32, 64, 128, 256)):
            if nx % v == 0:
                out.append(_hip.Tile(r, z, v, 4, 0))
    elif kind == "tb" and dims == 3:
        for (v, r, w), z in itertools.product(TB3_SHAPES, (0, 64, 96, 128, 192, 256)):
            if nx % v == 0:
                out.append(_hip.Tile(r, z, v, w, 0))
    elif kind == "tb":
        for v, z in itertools.product(vecs, (0, 64, 128, 256, 512)):
            if nx % v == 0:
                out.append(_hip.Tile(0, z, v, 4, 0))
    return out


def tune(key: str, shape: tuple[int, ...], bc: str = "periodic", steps: int = 20,
         verbose: bool = True, source: bool = False) -> dict[str, Any]:
    """Time every candidate tile of variant ``key`` on ``shape`` (with a float32 source when
    ``source``); return the best entry."""
    from .problem import heat_stencil
    from .registry import VARIANTS, make_stepper

    var = VARIANTS[key]
    dims = len(shape)
    u0 = np.full(shape, 0.5, np.float32)
    b = np.full(shape, 1e-5, np.float32) if source else None
    st = make_stepper(key, u0, heat_stencil(dims, 0.9 / (2 * dims), bc), b)
    kind, k = st.kind, st.steps_per_pass
    ev0, ev1 = _hip.Event(), _hip.Event()
    best = None
    try:
        for tile in candidates(kind, k, dims, shape[-1], var.storage):
            st.tile_override = tile
            try:
                # warm-up with the timed call itself: a small grid's stepper captures its graph
                # of passes here, so the timed call measures replays, as a user's run does
                st.step(steps * k)
                st.synchronize()
                ev0.record(st.stream.handle)
                st.step(steps * k)
                ev1.record(st.stream.handle)
                ev1.sync()
            except _hip.HipError as e:
                if verbose:
                    print(f"  {tile.as_dict()} failed: {e}")
                continue
            ms = ev0.elapsed_ms(ev1) / (steps * k)
            glups = int(np.prod(shape)) / ms / 1e6
            if verbose:
                print(f"  {key} {tile.as_dict()}: {glups:8.1f} GLUPS", flush=True)
            if best is None or glups > best["glups"]:
                best = {"tile": tile.as_dict(), "glups": glups, "shape": list(shape)}
    finally:
        ev0.close()
        ev1.close()
        st.close()
    if best is None:
        raise RuntimeError(f"no candidate tile ran for {key} {shape}")
    best["key"] = table_key(kind, k, dims, var.storage, bc, shape, source)
    return best


def update_table(entries: list[dict[str, Any]], path: Path = TABLE) -> None:
    global _table
    table = json.loads(path.read_text()) if path.exists() else {}
    info = _hip.device_info(0)
    for e in entries:
        table[e["key"]] = {"tile": e["tile"], "glups": round(e["glups"], 2), "shape": e["shape"],
                           "arch": info["arch"]}
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(table, indent=1, sort_keys=True))
    _table = None
