"""The benchmark entry point (reference ``bench/run.py``), for the ``hip-cdna4`` family.

    python -m pyheat3d_amd.bench.run --quick   # small sizes   -> results/quick.json
    python -m pyheat3d_amd.bench.run --full    # full sweep    -> results/mi355x.json
    python -m pyheat3d_amd.bench.run --dev     # dev sizes     -> results/dev.json

Optional filters: ``--variants tb zmarch`` (substring match), ``--dims 3``.
"""

from __future__ import annotations

import argparse

from ..benchmark import run


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    g = ap.add_mutually_exclusive_group()
    g.add_argument("--quick", action="store_true")
    g.add_argument("--full", action="store_true")
    g.add_argument("--dev", action="store_true")
    ap.add_argument("--variants", nargs="*", default=None)
    ap.add_argument("--dims", type=int, nargs="*", default=[2, 3])
    ap.add_argument("--out", default=None)
    ap.add_argument("--out-dir", default="results")
    a = ap.parse_args()
    run("full" if a.full else ("dev" if a.dev else "quick"), a.out_dir, a.variants,
        tuple(a.dims), a.out)


if __name__ == "__main__":
    main()
