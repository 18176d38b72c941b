"""Fixed launch sequence for hardware-counter runs under ``rocprofv3``.

    python -m pyheat3d_amd.profile_run --variant hip-cdna4/zmarch --n 1024 --steps 10

Runs ``--copies`` launches of the library's copy kernel on a buffer of ``--copy-bytes`` (the
known-size reference that scales FETCH_SIZE / WRITE_SIZE to bytes), then ``--steps`` sweeps of
the variant on an ``n^3`` periodic field (constant initial value: counters do not depend on the
data). Prints one line per phase so a slow or stuck run is visible.
"""

from __future__ import annotations

import argparse

import numpy as np

from . import _hip
from .problem import heat_stencil
from .registry import make_stepper


def summarize(fetch_csv: str, write_csv: str, match: str, copy_bytes: int,
              algorithmic_bytes: float) -> dict:
    """HBM bytes per launch of the kernel whose name contains ``match``, from two counter runs.

    Each counter is scaled by (known bytes / counter reading) of the copy kernel in the same
    run: the copy kernel reads and writes exactly ``copy_bytes``.
    """
    import csv

    def avg(path: str, counter: str) -> tuple[float, float]:
        kern, copy = [], []
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            v = float(r["Counter_Value"])
            if "copy16_kernel" in r["Kernel_Name"]:
                copy.append(v)
            elif match in r["Kernel_Name"]:
                kern.append(v)
        if not kern or not copy:
            raise SystemExit(f"{path}: no {counter} rows for '{match}' or the copy kernel")
        return sum(kern) / len(kern), sum(copy) / len(copy)

    fk, fc = avg(fetch_csv, "FETCH_SIZE")
    wk, wc = avg(write_csv, "WRITE_SIZE")
    kib = copy_bytes / 1024.0
    ff, wf = kib / fc, kib / wc
    traffic = (fk * ff + wk * wf) * 1024.0
    return {"kernel": match, "fetch_kib": fk, "write_kib": wk, "copy_bytes": copy_bytes,
            "copy_fetch_kib": fc, "copy_write_kib": wc, "fetch_scale": ff, "write_scale": wf,
            "hbm_bytes_per_launch": traffic, "algorithmic_bytes_per_launch": algorithmic_bytes,
            "traffic_over_algorithmic": traffic / algorithmic_bytes}


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--summarize", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarise two counter runs instead of running kernels")
    ap.add_argument("--match", default="zmarch3d<float, true, false, 4")
    ap.add_argument("--algorithmic-bytes", type=float, default=8.0 * 1024 ** 3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--variant", default="hip-cdna4/zmarch")
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--dims", type=int, default=3)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--copies", type=int, default=3)
    ap.add_argument("--tile", default=None, metavar="TY,ZCH,VEC,WAVES,FLAGS",
                    help="launch shape instead of the tuning table's")
    a = ap.parse_args()
    if a.summarize:
        import json

        out = {"variant": a.variant, "n": a.n,
               **summarize(*a.summarize, a.match, a.copy_bytes, a.algorithmic_bytes)}
        text = json.dumps(out, indent=1)
        print(text)
        if a.out:
            with open(a.out, "w") as f:
                f.write(text + "\n")
        return
    stream = _hip.Stream()
    src, dst = _hip.malloc(a.copy_bytes), _hip.malloc(a.copy_bytes)
    for _ in range(a.copies):
        _hip.check(_hip.lib().ph3_copy_kernel(src, dst, a.copy_bytes, stream.handle), "copy")
    stream.sync()
    _hip.free(src)
    _hip.free(dst)
    stream.close()
    print(f"copy kernel: {a.copies} x {a.copy_bytes} bytes", flush=True)
    shape = (a.n,) * a.dims
    st = make_stepper(a.variant, np.full(shape, 0.5, np.float32),
                      heat_stencil(a.dims, 0.15 if a.dims == 3 else 0.2, "periodic"))
    if a.tile:
        st.tile_override = _hip.Tile(*(int(v) for v in a.tile.split(",")))
    try:
        st.step(a.steps * st.steps_per_pass)
        st.synchronize()
    finally:
        st.close()
    print(f"{a.variant}: {a.steps} launches on {shape}, tile {st.tile}", flush=True)


if __name__ == "__main__":
    main()
