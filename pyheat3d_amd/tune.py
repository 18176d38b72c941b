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


# ---- batched solves -------------------------------------------------------------------------
# Two more kinds of entries tune the batch's tier-2 passes (ph3_batch_*), which the reference
# does not have (it times one grid at a time):
#   batch/tile/<d>d/b<bucket>     {"tile": id}: the tile shape of a solve's fused passes
#                                 (ph3_solve_desc.tile), a function of the solve alone;
#   batch/policy/<d>d/k<k>        {"target_tiles": T, "max_chunk": C}: how the (dims, kpass)
#                                 classes cut their grids into tiles (ph3_batch_params).
# Neither changes a result: tile shape and chunk never change a site's arithmetic.

def batch_tile(dims: int, shape: tuple[int, ...]) -> int:
    """Tuned tier-2 tile shape id of a solve (0: the library's rule)."""
    e = load_table().get(f"batch/tile/{dims}d/b{bucket(shape)}")
    return int(e["tile"]) if e else 0


def batch_policy() -> list[tuple[int, int, int, int]]:
    """(dims, kpass, target_tiles, max_chunk) of every tuned batch class policy."""
    out = []
    for key, e in load_table().items():
        if key.startswith("batch/policy/"):
            _, _, d, k = key.split("/")
            out.append((int(d[:-1]), int(k[1:]), int(e["target_tiles"]), int(e["max_chunk"])))
    return sorted(out)


#: the fused classes of the float32 batch (dims, kpass) and the grid sizes that land in them
#: (2D: 8-step passes from 4097 sites; 3D: 2-step passes below 128^3 sites, 4-step from there)
BATCH_CLASSES = {(2, 8): (65, 2048), (3, 2): (17, 127), (3, 4): (128, 256)}
#: (target_tiles, max_chunk) candidates per dims; (0, 0) is the library's default
BATCH_POLICIES = {2: ((0, 0), (1024, 256), (2048, 256), (4096, 256), (8192, 256), (2048, 128),
                      (4096, 128), (8192, 512)),
                  3: ((0, 0), (512, 128), (1024, 128), (2048, 128), (4096, 128), (8192, 128),
                      (2048, 256), (4096, 256))}
BATCH_SHAPES = {2: (1, 2, 3, 4, 5), 3: (1, 2, 3, 4, 5, 6)}   # tier2_shape ids per dims


def tuning_mix(n: int = 256, seed: int = 77, jacobi_cap: int = 2048, small: bool = False):
    """A mix the batch is tuned on: the mixed workload's size ranges (60% 2D N in 64..2048, 40%
    3D N in 32..256), half heat, half Jacobi with a short sweep cap; ``small``: the same with
    2D N up to 256 and 3D N up to 64 (many small grids)."""
    from .batch import make_mix

    return make_mix(n, seed, n2=(64, 256 if small else 2048), n3=(32, 64 if small else 256),
                    frac_jacobi=0.5, jacobi_cap=jacobi_cap)


def _class_of(req) -> tuple[int, int] | None:
    """The fused class a float32 request of the tuning mixes lands in (None: tier 1)."""
    if req.sites <= _hip.BATCH_SMALL_SITES:
        return None
    if req.dims == 2:
        return (2, 8)
    return (3, 4) if req.sites >= (1 << 21) else (3, 2)


def _half(req, c) -> int:
    """0 for the lower half of a class's size buckets, 1 for the upper half."""
    lo, hi = BATCH_CLASSES[c]
    b0, b1 = bucket((lo,) * c[0]), bucket((hi,) * c[0])
    return 1 if bucket(req.shape) > (b0 + b1) // 2 else 0


def tail_mix(seed: int = 91):
    """The solves that set the length of a full mixed run: Jacobi Laplace on the largest grids
    (2D N in 600..2048, 3D N in 128..256, two 2D solves per 3D one), which need 10^5-10^6
    sweeps; timed over a fixed number of rounds, not to completion."""
    from .batch import Request

    rng = np.random.default_rng(seed)
    out = []
    for i in range(48):
        dims = 3 if i % 3 == 2 else 2
        lo, hi = (600, 2048) if dims == 2 else (128, 256)
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        out.append(Request(shape=(n,) * dims, mode="jacobi", steps=10 ** 6,
                           seed=int(rng.integers(1, 2 ** 62)), rid=i))
    return out


def _time_mixes(policy: dict, shapes: dict, n: int, reps: int = 2,
                tail_rounds: int = 3) -> tuple[float, float, float]:
    """Seconds of the full-range and small-grid tuning mixes run to completion (best of
    ``reps``) and of ``tail_rounds`` rounds of the tail mix (after one untimed round)."""
    from .batch import BatchSolver

    def shaped(reqs):
        for r in reqs:
            c = _class_of(r)
            r.tile = shapes.get((c, _half(r, c)), 0) if c else 0   # 0: the library's rule
        return reqs

    params = [(d, k, t, ch) for (d, k), (t, ch) in policy.items()]
    out = []
    for small in (False, True):
        best = None
        for _ in range(reps):
            b = BatchSolver(shaped(tuning_mix(n, small=small)), round_sweeps=256, params=params,
                            tuned_tiles=False)
            try:
                sec = b.run()["seconds"]
            finally:
                b.close()
            best = sec if best is None else min(best, sec)
        out.append(best)
    b = BatchSolver(shaped(tail_mix()), round_sweeps=256, params=params, tuned_tiles=False,
                    compact=False)
    try:
        b.run(max_rounds=1)
        out.append(b.run(max_rounds=tail_rounds)["seconds"])
    finally:
        b.close()
    return out[0], out[1], out[2]


def tune_batch(n: int = 256, verbose: bool = True) -> list[dict[str, Any]]:
    """Autotune the batch's tier-2 passes on three mixes (:func:`tuning_mix`, full size ranges
    and small grids, run to completion; :func:`tail_mix`, the largest Jacobi solves, a few
    rounds): coordinate descent over each fused class's chunk policy (target tiles per class,
    longest chunk) and the tile shape of the lower and upper half of its size range. The score
    is the sum of the three mixes' times relative to the library's defaults. Returns table
    entries for ``update_batch_table``."""
    policy = {c: (0, 0) for c in BATCH_CLASSES}
    shapes: dict = {(c, h): 0 for c in BATCH_CLASSES for h in (0, 1)}
    base = _time_mixes(policy, shapes, n)
    score = lambda t: sum(x / y for x, y in zip(t, base))   # noqa: E731
    best_t, best = base, float(len(base))
    if verbose:
        print(f"  batch tuning mixes ({n} solves), library defaults: {base[0]:.3f} s (full), "
              f"{base[1]:.3f} s (small), {base[2]:.3f} s (tail rounds)", flush=True)
    for sweep in range(2):
        for c in BATCH_CLASSES:
            for cand in BATCH_POLICIES[c[0]]:
                if cand == policy[c]:
                    continue
                trial = {**policy, c: cand}
                t = _time_mixes(trial, shapes, n)
                if verbose:
                    print(f"  sweep {sweep} class {c} policy {cand}: {t[0]:.3f} / {t[1]:.3f} / "
                          f"{t[2]:.3f} s", flush=True)
                if score(t) < best:
                    best, best_t, policy = score(t), t, trial
            for h in (0, 1):
                for sid in BATCH_SHAPES[c[0]]:
                    if sid == shapes[(c, h)]:
                        continue
                    trial = {**shapes, (c, h): sid}
                    t = _time_mixes(policy, trial, n)
                    if verbose:
                        print(f"  sweep {sweep} class {c} half {h} shape {sid}: {t[0]:.3f} / "
                              f"{t[1]:.3f} / {t[2]:.3f} s", flush=True)
                    if score(t) < best:
                        best, best_t, shapes = score(t), t, trial
    if verbose:
        print(f"  tuned: {best_t[0]:.3f} s (full), {best_t[1]:.3f} s (small), {best_t[2]:.3f} s "
              f"(tail rounds)", flush=True)
    entries = []
    for c, (lo, hi) in BATCH_CLASSES.items():
        d, k = c
        t, ch = policy[c]
        entries.append({"key": f"batch/policy/{d}d/k{k}", "target_tiles": t, "max_chunk": ch,
                        "mix_seconds": [round(x, 4) for x in best_t],
                        "default_mix_seconds": [round(x, 4) for x in base], "mix_solves": n})
        b0, b1 = bucket((lo,) * d), bucket((hi,) * d)
        for bk in range(b0, b1 + 1):
            h = 1 if bk > (b0 + b1) // 2 else 0
            entries.append({"key": f"batch/tile/{d}d/b{bk}", "tile": shapes[(c, h)]})
    return entries


def update_batch_table(entries: list[dict[str, Any]], path: Path = TABLE) -> None:
    """Write batch entries (``tune_batch``) into the tuning table."""
    global _table
    table = json.loads(path.read_text()) if path.exists() else {}
    arch = _hip.device_info(0)["arch"]
    for e in entries:
        table[e["key"]] = {**{k: v for k, v in e.items() if k != "key"}, "arch": arch}
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text(json.dumps(table, indent=1, sort_keys=True))
    _table = None
