"""Command line: ``python -m pyheat3d_amd {list, run, validate, bench, batch, serve, jacobi, tune}``.

``list``, ``run``, ``validate`` and ``bench`` mirror the reference CLI (``src/pyheat3d/cli.py``);
``batch``, ``serve``, ``jacobi`` and ``tune`` drive the batched solver, the serving loop, the
Jacobi-to-tolerance solve of the manufactured Poisson problem and the autotuner.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import analytic as an
from .problem import heat_stencil, jacobi_stencil, stability_limit
from .registry import VARIANTS, availability, make_stepper


def _cmd_list(_: argparse.Namespace) -> int:
    """Variants grouped by rung, with what each needs and whether this machine can run it."""
    rows = []
    for key, v in sorted(VARIANTS.items(), key=lambda kv: (kv[1].rung, kv[0])):
        ok, why = availability(key)
        rows.append((str(v.rung), key, "/".join(f"{d}D" for d in v.dims), v.storage,
                     str(v.steps_per_pass), "yes" if v.source else "no",
                     "ready" if ok else f"unavailable: {why}"))
    head = ("rung", "key", "dims", "storage", "k", "source", "here")
    width = [max(len(r[c]) for r in rows + [head]) for c in range(len(head) - 1)]
    for r in [head] + rows:
        print("  ".join(x.ljust(w) for x, w in zip(r, width)) + "  " + r[-1])
    return 0


def _initial_state(a: argparse.Namespace):
    """Stencil and float32 initial field of a ``run`` invocation, plus the per-step decay of
    the sine mode when that is the initial state (for the closed-form check)."""
    shape = (a.n,) * a.dims
    if a.mode == "jacobi":
        u0 = an.harmonic_sinh(a.n, a.dims).astype(np.float32)
        u0[(slice(1, -1),) * a.dims] = 0.0      # Laplace: boundary data, zero guess
        return jacobi_stencil(a.dims, a.bc), u0, None
    r = a.r if a.r is not None else 0.9 * stability_limit(a.dims)
    if a.init == "random":
        return heat_stencil(a.dims, r, a.bc), np.random.default_rng(0).random(shape).astype(np.float32), None
    g = an.ftcs_amplification(r, a.n, a.dims, a.bc)
    return heat_stencil(a.dims, r, a.bc), an.fourier_mode(a.n, a.dims, a.bc).astype(np.float32), g


def _cmd_run(a: argparse.Namespace) -> int:
    """One timed run of one variant; with the sine-mode initial state the final amplitude is
    checked against the closed form g^steps."""
    stencil, u0, g = _initial_state(a)
    st = make_stepper(a.variant, u0, stencil)
    try:
        st.step(st.steps_per_pass)           # compile-free warm-up: one pass
        st.synchronize()
        t0 = time.perf_counter()
        st.step(a.steps)
        st.synchronize()
        dt = time.perf_counter() - t0
        u = st.get()
        sites = st.ncells
        k = st.steps_per_pass
    finally:
        st.close()
    report = {"variant": a.variant, "shape": list(u0.shape), "bc": a.bc, "mode": a.mode,
              "steps_timed": a.steps, "steps_per_pass": k, "seconds": round(dt, 6),
              "GLUPS": round(sites * a.steps / dt / 1e9, 3),
              "field": {"min": float(u.min()), "max": float(u.max()), "mean": float(u.mean())}}
    if g is not None:
        total = a.steps + k
        ref = u0.astype(np.float64) * g ** total
        report["mode_check"] = {"steps": total, "closed_form_amplitude": g ** total,
                                "max_rel_dev": float(np.abs(u - ref).max() / max(np.abs(ref).max(), 1e-300))}
    for key, val in report.items():
        print(f"{key:>14}: {val}")
    print("(one run after one warm-up pass: a quick check, not the benchmark protocol; see bench)")
    return 0


def _cmd_validate(a: argparse.Namespace) -> int:
    from .validation import run_all

    res = run_all(quick=a.quick)
    bad = 0
    for x in res["spatial"]:
        ok = 1.85 < x["order"] < 2.15
        bad += not ok
        print(f"spatial   {x['variant']:22s} {x['dims']}D order {x['order']:.3f} {'ok' if ok else 'FAIL'}")
    for x in res["temporal"]:
        ok = 0.95 < x["order"] < 1.05
        bad += not ok
        print(f"temporal  {x['variant']:22s} {x['dims']}D order {x['order']:.3f} {'ok' if ok else 'FAIL'}")
    for name in ("laplace", "poisson"):
        for x in res[name]:
            ok = 1.85 < x["order"] < 2.15
            bad += not ok
            print(f"{name:9s} {x['variant']:22s} {x['dims']}D order {x['order']:.3f} {'ok' if ok else 'FAIL'}")
    for x in res["laplace_quadratic"]:
        print(f"quadratic {x['variant']:22s} {x['dims']}D n={x['n']} {x['status']} after "
              f"{x['sweeps']} sweeps, err {x['err']:.2e}")
    for x in res["poisson_eigen"]:
        print(f"poisson   {x['variant']:22s} n={x['n']} {x['status']} sweeps {x['sweeps']} "
              f"(predicted {x['predicted_sweeps']}), err vs FP64 {x['err_vs_fp64']:.2e}")
    if a.out:
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        Path(a.out).write_text(json.dumps(res, indent=1))
    return 1 if bad else 0


def _cmd_bench(a: argparse.Namespace) -> int:
    from .benchmark import run

    run("full" if a.full else ("dev" if a.dev else "quick"), a.out_dir, a.variants, tuple(a.dims))
    return 0


def _cmd_batch(a: argparse.Namespace) -> int:
    from .benchmark import batch_benchmark

    r = batch_benchmark(a.n, a.seed, (a.n2_min, a.n2_max), (a.n3_min, a.n3_max),
                        a.frac_jacobi, a.storage, a.baseline_every, a.round_sweeps,
                        a.jacobi_cap, a.accuracy_every, use_graph=not a.no_graph,
                        graphs_baseline=not a.no_graphs_baseline, progress=True)
    print(json.dumps(r, indent=1))
    return 0


def _cmd_serve(a: argparse.Namespace) -> int:
    from .benchmark import serve_benchmark

    r = serve_benchmark(a.n, a.rate, a.seed, (a.n2_min, a.n2_max), (a.n3_min, a.n3_max),
                        a.storage, a.budget_gb, a.round_sweeps, a.jacobi_cap, a.accuracy_every,
                        use_graph=not a.no_graph, progress=True)
    print(json.dumps(r, indent=1))
    return 0


def _cmd_jacobi(a: argparse.Namespace) -> int:
    from .validation import poisson_eigen

    last = [time.perf_counter()]

    def progress(sweeps: int, rel: float) -> None:
        if time.perf_counter() - last[0] > 10.0:
            last[0] = time.perf_counter()
            print(f"{a.variant} n={a.n}: {sweeps} sweeps, rel residual {rel:.3e}",
                  file=sys.stderr, flush=True)

    r = poisson_eigen(a.variant, a.n, 3, a.tol, a.check_every, a.max_sweeps, progress)
    print(json.dumps(r, indent=1))
    return 0


def _cmd_tune(a: argparse.Namespace) -> int:
    from .tune import tune, tune_batch, update_batch_table, update_table

    if a.batch:
        entries = tune_batch(a.batch_n)
        for e in entries:
            print(f"best {e['key']}: " + ", ".join(f"{k}={v}" for k, v in e.items() if k != "key"))
        if a.write:
            update_batch_table(entries)
        return 0

    entries = []
    for key in a.variants:
        dims = a.dims if a.dims in VARIANTS[key].dims else VARIANTS[key].dims[0]
        e = tune(key, (a.n,) * dims, a.bc, source=a.source)
        print(f"best {key}: {e['tile']} {e['glups']:.1f} GLUPS")
        entries.append(e)
    if a.write:
        update_table(entries)
    return 0


def _size_args(p: argparse.ArgumentParser) -> None:
    p.add_argument("--n2-min", type=int, default=64)
    p.add_argument("--n2-max", type=int, default=2048)
    p.add_argument("--n3-min", type=int, default=32)
    p.add_argument("--n3-max", type=int, default=256)
    p.add_argument("--jacobi-cap", type=int, default=10 ** 6, help="sweep cap of Jacobi solves")
    p.add_argument("--no-graph", action="store_true",
                   help="enqueue rounds eagerly instead of replaying their HIP graphs (for kernel "
                        "traces: graph launches of large rounds have crashed under rocprofv3)")
    p.add_argument("--accuracy-every", type=int, default=8,
                   help="compare every k-th result with its FP64 NumPy solution (bounded size)")
    p.add_argument("--storage", default="float32", choices=("float32", "float16", "bfloat16"))
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--round-sweeps", type=int, default=256)


def main(argv: list[str] | None = None) -> int:
    p = argparse.ArgumentParser(prog="pyheat3d_amd", description=__doc__)
    sub = p.add_subparsers(dest="cmd", required=True)
    sub.add_parser("list", help="list variants and availability").set_defaults(fn=_cmd_list)

    r = sub.add_parser("run", help="run one variant once and print a rough rate")
    r.add_argument("--variant", default="hip-cdna4/naive", choices=sorted(VARIANTS))
    r.add_argument("--dims", type=int, default=3, choices=(2, 3))
    r.add_argument("--n", type=int, default=64)
    r.add_argument("--steps", type=int, default=100)
    r.add_argument("--bc", default="dirichlet", choices=("dirichlet", "periodic"))
    r.add_argument("--mode", default="heat", choices=("heat", "jacobi"))
    r.add_argument("--r", type=float, default=None)
    r.add_argument("--init", default="mode", choices=("mode", "random"))
    r.set_defaults(fn=_cmd_run)

    v = sub.add_parser("validate", help="convergence studies against closed forms")
    v.add_argument("--quick", action="store_true")
    v.add_argument("--out", default=None)
    v.set_defaults(fn=_cmd_validate)

    b = sub.add_parser("bench", help="single-grid benchmark sweep (as bench/run.py)")
    g = b.add_mutually_exclusive_group()
    g.add_argument("--quick", action="store_true")
    g.add_argument("--full", action="store_true")
    g.add_argument("--dev", action="store_true")
    b.add_argument("--variants", nargs="*", default=None)
    b.add_argument("--dims", type=int, nargs="*", default=[2, 3])
    b.add_argument("--out-dir", default="results")
    b.set_defaults(fn=_cmd_bench)

    ba = sub.add_parser("batch", help="batched solves vs one launch per solve")
    ba.add_argument("--n", type=int, default=2048)
    ba.add_argument("--frac-jacobi", type=float, default=0.5)
    ba.add_argument("--baseline-every", type=int, default=8)
    ba.add_argument("--no-graphs-baseline", action="store_true",
                    help="skip the per-solve leg on the stepper's replayed graphs")
    _size_args(ba)
    ba.set_defaults(fn=_cmd_batch)

    se = sub.add_parser("serve", help="continuous-batching server on Poisson arrivals")
    se.add_argument("--n", type=int, default=1024)
    se.add_argument("--rate", type=float, default=500.0)
    se.add_argument("--budget-gb", type=float, default=200.0)
    _size_args(se)
    se.set_defaults(fn=_cmd_serve)

    j = sub.add_parser("jacobi", help="Jacobi to tolerance on the manufactured Poisson problem")
    j.add_argument("--variant", default="hip-cdna4/zmarch", choices=sorted(VARIANTS))
    j.add_argument("--n", type=int, default=768)
    j.add_argument("--tol", type=float, default=1e-6)
    j.add_argument("--check-every", type=int, default=500)
    j.add_argument("--max-sweeps", type=int, default=3 * 10 ** 6)
    j.set_defaults(fn=_cmd_jacobi)

    t = sub.add_parser("tune", help="autotune launch shapes and update the tuning table")
    t.add_argument("--variants", nargs="+", default=["hip-cdna4/zmarch", "hip-cdna4/tb3"])
    t.add_argument("--dims", type=int, default=3)
    t.add_argument("--n", type=int, default=1024)
    t.add_argument("--bc", default="periodic", choices=("dirichlet", "periodic"))
    t.add_argument("--source", action="store_true", help="tune with a float32 source term")
    t.add_argument("--write", action="store_true")
    t.add_argument("--batch", action="store_true",
                   help="tune the batch's tier-2 passes (tile shapes, chunk policies) instead")
    t.add_argument("--batch-n", type=int, default=256, help="solves in the tuning mix")
    t.set_defaults(fn=_cmd_tune)

    a = p.parse_args(argv)
    return a.fn(a)


if __name__ == "__main__":
    sys.exit(main())
