"""Autotuner entry point (reference ``bench/tune.py``): sweep launch shapes, update the table.

    python -m pyheat3d_amd.bench.tune hip-cdna4/zmarch hip-cdna4/tb3 --n 1024 --write
"""

from __future__ import annotations

import argparse

from ..tune import tune, update_table


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("variants", nargs="+")
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--dims", type=int, default=3)
    ap.add_argument("--bc", default="periodic")
    ap.add_argument("--write", action="store_true")
    a = ap.parse_args()
    entries = [tune(k, (a.n,) * a.dims, a.bc) for k in a.variants]
    if a.write:
        update_table(entries)


if __name__ == "__main__":
    main()
