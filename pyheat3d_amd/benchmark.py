"""Benchmark driver behind ``python -m pyheat3d_amd bench`` (reference ``src/pyheat3d/benchmark.py``).

Single-grid timing keeps the reference's protocol and record schema (``benchmark.py:145-234``):
warm-up of ``2k`` steps, calibration until a repeat lasts >= 0.2 ``min_time``, ``repeats`` timed
repeats of ``n`` steps between device synchronisations, median MLUPS over interior sites,
minimal-model bandwidth. Ceilings are MI355X ones: the copy bandwidth measured in the same
process by this library's copy kernel, and the 8 TB/s spec.

:func:`batch_benchmark` measures the batched server path against one launch per solve on the
same request mix; :func:`serve_benchmark` runs the continuous-batching server on Poisson arrivals.
"""

from __future__ import annotations

import datetime as _dt
import json
import math
import os
import platform
import time
from pathlib import Path
from typing import Any

import numpy as np

from . import traffic
from .problem import heat_stencil
from .registry import VARIANTS, availability, hip_available, make_stepper

PLANS: dict[str, dict[str, Any]] = {
    "quick": {"sizes": {2: [1024, 4096], 3: [64, 256]}, "min_time": 0.05, "repeats": 5,
              "mem_limit_GB": 8.0, "out": "quick.json"},
    "dev": {"sizes": {2: [512, 1024, 2048, 4096, 8192], 3: [32, 64, 128, 256, 320]},
            "min_time": 0.1, "repeats": 5, "mem_limit_GB": 8.0, "out": "dev.json"},
    "full": {"sizes": {2: [1024, 2048, 4096, 8192, 16384, 32768], 3: [64, 128, 256, 512, 1024]},
             "min_time": 0.5, "repeats": 5, "mem_limit_GB": 64.0, "out": "mi355x.json"},
}


_SMI: Any = None


def _smi_handle():
    """AMD SMI handle of GPU 0, or None (library absent or no access)."""
    global _SMI
    if _SMI is None:
        try:
            import amdsmi

            amdsmi.amdsmi_init()
            _SMI = (amdsmi, amdsmi.amdsmi_get_processor_handles()[0])
        except Exception:
            _SMI = False
    return _SMI or None


def gpu_state() -> dict[str, Any]:
    """GPU clocks, temperature and power right now (AMD SMI; the reference reads the same
    through NVML for its throttling audit, ``benchmark.py:86-100``). Empty when unavailable."""
    smi = _smi_handle()
    if smi is None:
        return {}
    amdsmi, h = smi
    out: dict[str, Any] = {}

    def clk(kind):
        info = amdsmi.amdsmi_get_clock_info(h, kind)
        return info.get("clk", info.get("cur_clk"))

    for key, fn in (("gfx_clock_MHz", lambda: clk(amdsmi.AmdSmiClkType.GFX)),
                    ("mem_clock_MHz", lambda: clk(amdsmi.AmdSmiClkType.MEM)),
                    ("temp_hotspot_C", lambda: amdsmi.amdsmi_get_temp_metric(
                        h, amdsmi.AmdSmiTemperatureType.HOTSPOT,
                        amdsmi.AmdSmiTemperatureMetric.CURRENT)),
                    ("power_W", lambda: (lambda p: p.get("current_socket_power",
                                                         p.get("average_socket_power")))(
                        amdsmi.amdsmi_get_power_info(h))),
                    ("gfx_activity_pct", lambda: amdsmi.amdsmi_get_gpu_activity(h).get("gfx_activity"))):
        try:
            v = fn()
            if v is not None and v != "N/A":
                out[key] = v
        except Exception:
            pass
    return out


def cpu_model() -> str | None:
    """The host CPU's model name (``lscpu``'s "Model name", read from /proc/cpuinfo)."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or None


def machine_info() -> dict[str, Any]:
    from . import _hip

    info: dict[str, Any] = {"cpu": cpu_model(), "cpu_count": os.cpu_count(),
                            "python": platform.python_version(), "numpy": np.__version__}
    try:
        d = _hip.device_info(0)
        info.update(gpu=d["name"] or "MI355X", arch=d["arch"], cu_count=d["cu_count"],
                    hbm_bytes=d["total_mem_bytes"])
        ver = Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / ".info" / "version"
        info["driver"] = "ROCm " + ver.read_text().strip() if ver.exists() else None
    except Exception as e:  # no device
        info["gpu"] = f"unavailable: {e}"
    info["date"] = _dt.datetime.now().astimezone().isoformat(timespec="seconds")
    return info


def time_variant(key: str, u0: np.ndarray, min_time: float, repeats: int,
                 r: float | None = None, copy_bw: float | None = None,
                 bc: str = "dirichlet") -> dict[str, Any]:
    """Time one variant on one initial state; returns a reference-schema record."""
    var = VARIANTS[key]
    dims = u0.ndim
    st = make_stepper(key, u0, heat_stencil(dims, r if r is not None else 0.9 / (2 * dims), bc))
    k = var.steps_per_pass
    tile = dict(st.tile)
    try:
        t0 = time.perf_counter()
        st.step(2 * k)
        st.synchronize()
        warm = time.perf_counter() - t0
        n = k
        while True:
            t0 = time.perf_counter()
            st.step(n)
            st.synchronize()
            dt = time.perf_counter() - t0
            if dt >= 0.2 * min_time or n >= 1 << 22:
                break
            n *= 2
        n = max(k, int(math.ceil(min_time / (dt / n) / k)) * k)
        times = []
        for _ in range(repeats):
            st.synchronize()
            t0 = time.perf_counter()
            st.step(n)
            st.synchronize()
            times.append(time.perf_counter() - t0)
        finite = bool(np.isfinite(st.get()).all())
    finally:
        st.close()
    ncells = math.prod(s - 2 for s in u0.shape) if bc == "dirichlet" else math.prod(u0.shape)
    mlups = [ncells * n / t / 1e6 for t in times]
    med = float(np.median(mlups))
    bpu = traffic.minimal_bytes(var.bytes_per_value, k)
    eff = med * bpu / 1e3
    rec: dict[str, Any] = {
        "backend": var.backend, "variant": var.name, "key": key, "rung": var.rung, "dims": dims,
        "size": list(u0.shape), "bc": bc, "dtype": var.storage, "metric": "MLUPS", "value": med,
        "repeats": repeats, "median": med, "min": float(min(mlups)), "max": float(max(mlups)),
        "steps_per_repeat": n, "seconds": [float(t) for t in times], "steps_per_pass": k,
        "bytes_per_unit": bpu, "effective_GBs": eff,
        "pct_copy_bw": 100.0 * eff / copy_bw if copy_bw else None,
        "pct_spec_bw": 100.0 * eff / traffic.SPEC_BW_GBS,
        "roofline_MLUPS": traffic.roofline_mlups(copy_bw or traffic.SPEC_BW_GBS, bpu),
        "flops_per_unit": traffic.flops_per_update(dims), "GFLOPs": med * traffic.flops_per_update(dims) / 1e3,
        "warmup_s": warm, "finite": finite, "gpu_state": gpu_state(), "tile": tile,
        "notes": var.description,
    }
    if k > 1:
        rec["single_step_equiv_GBs"] = med * traffic.minimal_bytes(var.bytes_per_value, 1) / 1e3
        hb = traffic.halo_model_bytes(dims, k, var.bytes_per_value, tile)
        rec["halo_model_bytes_per_unit"] = hb
        rec["halo_model_GBs"] = med * hb / 1e3
    return rec


def run(mode: str = "quick", out_dir: str | Path = "results", variants: list[str] | None = None,
        dims_list: tuple[int, ...] = (2, 3), out_name: str | None = None,
        verbose: bool = True) -> dict[str, Any]:
    """Run a plan over sizes x variants and write the results JSON (reference ``run``)."""
    from .roofline import copy_bandwidth

    plan = PLANS[mode]
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    out_path = out_dir / (out_name or plan["out"])
    keys = [k for k in VARIANTS if variants is None or any(v in k for v in variants)]
    keys = [k for k in keys if availability(k)[0]]
    copy_bw = copy_bandwidth(1 << 30) if hip_available() else None
    result: dict[str, Any] = {
        "machine": machine_info(), "mode": mode,
        # the reference's figure script reads copy_bandwidth_GBs (its roofline lines) and
        # measured_copy_bandwidth_GBs (figures/make_figures.py:306-310)
        "ceiling": {"copy_bandwidth_GBs": copy_bw or traffic.COPY_BW_GBS,
                    "measured_copy_bandwidth_GBs": copy_bw,
                    "spec_bandwidth_GBs": traffic.SPEC_BW_GBS},
        "traffic_model": "minimal: 2*bytes_per_value/steps_per_pass per lattice-site update",
        "runs": [],
    }
    t_start = time.perf_counter()
    rng = np.random.default_rng(12345)
    for dims in dims_list:
        for n in plan["sizes"][dims]:
            shape = (n,) * dims
            u0 = rng.random(shape, dtype=np.float32)
            for key in keys:
                var = VARIANTS[key]
                if dims not in var.dims or 2 * math.prod(shape) * var.bytes_per_value / 1e9 > plan["mem_limit_GB"]:
                    continue
                try:
                    rec = time_variant(key, u0, plan["min_time"], plan["repeats"], copy_bw=copy_bw)
                except Exception as e:  # keep sweeping, record the failure
                    print(f"FAILED {key} {shape}: {type(e).__name__}: {e}")
                    continue
                result["runs"].append(rec)
                if verbose:
                    print(f"{dims}D {n:>6} {key:26s} {rec['median'] / 1e3:8.2f} GLUPS "
                          f"{rec['effective_GBs']:8.1f} GB/s", flush=True)
            out_path.write_text(json.dumps(result, indent=1))
    result["wall_time_s"] = time.perf_counter() - t_start
    out_path.write_text(json.dumps(result, indent=1))
    return result


# ----------------------------------------------------------------------------------------------
# batched and served solves
# ----------------------------------------------------------------------------------------------

def _lat_stats(ms) -> dict[str, float | None]:
    ms = np.asarray(ms, dtype=np.float64)
    if ms.size == 0:
        return {"p50_ms": None, "p99_ms": None, "max_ms": None}
    return {"p50_ms": float(np.percentile(ms, 50)), "p99_ms": float(np.percentile(ms, 99)),
            "max_ms": float(ms.max())}


def reached(req, status: str) -> bool:
    """A heat solve that ran its steps, or a Jacobi solve whose FP64 relative residual reached
    its tolerance (not one that stagnated at the storage floor or hit the sweep cap)."""
    return status == "done"


def outcome_stats(reqs, results, seconds: float) -> dict[str, Any]:
    """Solves/s and completion-latency percentiles, over every solve and per outcome: reached
    the target (heat done, Jacobi at tol), stagnated at the storage floor, capped."""
    groups: dict[str, list] = {"reached_target": [], "stagnated": [], "max_sweeps": []}
    for q, r in zip(reqs, results):
        key = "reached_target" if reached(q, r.status) else r.status
        groups.setdefault(key, []).append(r)
    out: dict[str, Any] = {"solves": len(results), "seconds": seconds,
                           "solves_per_s": len(results) / seconds,
                           **_lat_stats([r.seconds * 1e3 for r in results])}
    for key, rs in groups.items():
        out[key] = {"solves": len(rs), "solves_per_s": len(rs) / seconds,
                    **_lat_stats([r.seconds * 1e3 for r in rs])}
    jac = [(q, r) for q, r in zip(reqs, results) if q.mode == "jacobi"]
    out["jacobi"] = {"solves": len(jac),
                     "converged": sum(r.status == "done" for _, r in jac),
                     "stagnated": sum(r.status == "stagnated" for _, r in jac),
                     "capped": sum(r.status == "max_sweeps" for _, r in jac),
                     "max_rel_residual": max((r.rel_residual or 0.0) for _, r in jac) if jac else None}
    return out


def accuracy_sample(reqs, max_sites: int = 1 << 22, max_heat_work: float = 2e9,
                    every: int = 8) -> set[int]:
    """Request ids whose error against the FP64 solution is measured: every ``every``-th request
    whose FP64 solution costs seconds on one core (Jacobi: at most ``max_sites`` sites, one DST
    solve; heat: at most ``max_heat_work`` site-steps in float64)."""
    out = set()
    for q in reqs[::every]:
        if q.mode == "heat" and q.sites * q.steps <= max_heat_work:
            out.add(q.rid)
        elif q.mode == "jacobi" and q.sites <= max_sites:
            out.add(q.rid)
    return out


def accuracy_stats(reqs, results, progress: bool = False) -> dict[str, Any]:
    """Max and median error vs FP64 per (mode, outcome) over the results that carry a field."""
    import time as _time

    from .batch import error_vs_fp64

    rows = []
    t_last = _time.perf_counter()
    for q, r in zip(reqs, results):
        if r.u is None:
            continue
        if progress and _time.perf_counter() - t_last > 30:
            t_last = _time.perf_counter()
            _printer("accuracy: {0} results compared with FP64 solutions")(len(rows))
        e = error_vs_fp64(q, r.u)
        r.extra["err_vs_fp64"] = e
        r.u = None
        rows.append((q.mode, "reached_target" if reached(q, r.status) else r.status, e))
    out: dict[str, Any] = {"measured": len(rows)}
    for mode in ("heat", "jacobi"):
        for oc in ("reached_target", "stagnated", "max_sweeps"):
            es = [e for m, o, e in rows if m == mode and o == oc]
            if es:
                out[f"{mode}/{oc}"] = {"n": len(es), "max": float(max(es)),
                                       "median": float(np.median(es))}
    return out


def batch_benchmark(n: int = 2048, seed: int = 2024, n2: tuple[int, int] = (64, 2048),
                    n3: tuple[int, int] = (32, 256), frac_jacobi: float = 0.5,
                    storage: str = "float32", baseline_every: int = 8, round_sweeps: int = 256,
                    jacobi_cap: int = 10 ** 6, accuracy_every: int = 8,
                    graphs_baseline: bool = True, use_graph: bool = True,
                    progress: bool = False) -> dict[str, Any]:
    """All requests submitted at t = 0: batched graph rounds vs one launch per pass per solve.

    The per-solve leg runs every ``baseline_every``-th request (a systematic sample of the same
    mix), each with the fused-step rung the batch chose for it (the same float32 bits), so the
    speedup measures batching alone; its solves/s is the sample size over the sum of its solve
    times. With ``graphs_baseline`` the sample also runs on the single-grid stepper's replayed
    graphs (the stronger per-solve path for small grids). Solves/s and completion latency are
    reported over all solves and separately for those that reached their target (heat: all
    steps; Jacobi: the FP64 relative residual at ``tol``), stagnated at the storage floor or hit
    the sweep cap; a sample of final fields is compared with FP64 NumPy solutions.
    ``use_graph=False`` enqueues every round eagerly instead of replaying its captured HIP graph
    (for kernel traces: see the known limits of the batch).
    """
    from .batch import BatchSolver, batch_kpass, make_mix, solve_one_by_one

    reqs = make_mix(n, seed, n2=n2, n3=n3, frac_jacobi=frac_jacobi, storage=storage,
                    jacobi_cap=jacobi_cap)
    acc_ids = accuracy_sample(reqs, every=accuracy_every)
    bs = BatchSolver(reqs, round_sweeps=round_sweeps, use_graph=use_graph)
    try:
        timing = bs.run(progress=_printer("batch: round {0}, {1} running, {2:.0f} s")
                        if progress else None)
        kp = batch_kpass(bs)
        res = bs.results(keep_fields=acc_ids)
    finally:
        bs.close()
    work = sum(q.sites * r.steps for q, r in zip(reqs, res))
    sample = reqs[::baseline_every]
    if progress:
        _printer("batch: {0} solves in {1:.1f} s; per-solve leg on {2} of them")(
            n, timing["seconds"], len(sample))
    show = _printer("per-solve: {0}/{1} solves, {2:.0f} s") if progress else None
    base, _ = solve_one_by_one(sample, kpass=kp, progress=show)
    base_time = sum(r.seconds for r in base)
    by_rid = {r.rid: r for r in res}
    agree = sum(1 for b in base if by_rid[b.rid].steps == b.steps and by_rid[b.rid].status == b.status)
    out: dict[str, Any] = {
        "workload": f"{n} solves, {int(round(100 * (1 - frac_jacobi)))}% heat (200 steps) / "
                    f"{int(round(100 * frac_jacobi))}% Jacobi Laplace to 1e-6 (sweep cap "
                    f"{jacobi_cap}), 60% 2D N in {list(n2)}, 40% 3D N in {list(n3)}, {storage}, "
                    f"seed {seed}",
        "batched": {"rounds": timing["rounds"], "site_updates": work,
                    "effective_glups": work / timing["seconds"] / 1e9,
                    **outcome_stats(reqs, res, timing["seconds"])},
        "per_solve": {"sample": len(sample), "every": baseline_every,
                      "rung": "the batch's fused-step rung per solve (tb<kpass>, or the march)",
                      **outcome_stats(sample, base, base_time)},
        "speedup": (n / timing["seconds"]) / (len(sample) / base_time),
        "baseline_agrees_on_steps_and_status": f"{agree}/{len(sample)}",
        "sample_sweeps": {int(b.rid): int(by_rid[b.rid].steps) for b in base},
        "accuracy_vs_fp64": accuracy_stats(reqs, res, progress),
    }
    rs = out["batched"]["reached_target"]
    ps = out["per_solve"]["reached_target"]
    if ps["solves"] and rs["solves"]:
        out["speedup_reached_target"] = rs["solves_per_s"] / ps["solves_per_s"]
    if graphs_baseline:
        based, _ = solve_one_by_one(sample, graphs=True, kpass=kp, progress=show)
        based_time = sum(r.seconds for r in based)
        out["per_solve_graphs"] = {"seconds": based_time, "solves_per_s": len(sample) / based_time}
        out["speedup_vs_graphs"] = (n / timing["seconds"]) / (len(sample) / based_time)
    return out


def _printer(fmt: str):
    """A progress callback that prints one formatted line to stderr."""
    import sys

    def show(*args):
        print(fmt.format(*args), file=sys.stderr, flush=True)
    return show


def serve_benchmark(n: int = 1024, rate: float = 500.0, seed: int = 2024,
                    n2: tuple[int, int] = (64, 2048), n3: tuple[int, int] = (32, 256),
                    storage: str = "float32", budget_gb: float = 200.0,
                    round_sweeps: int = 128, jacobi_cap: int = 10 ** 6,
                    accuracy_every: int = 8, time_limit: float = 3600.0,
                    use_graph: bool = True, progress: bool = False) -> dict[str, Any]:
    """The mix arriving as a Poisson process of ``rate`` requests/s, served by continuous
    batching under an HBM budget; latency = completion - arrival, reported over all solves and
    per outcome, with a sample of results compared with FP64 NumPy solutions."""
    from .batch import make_mix
    from .server import Server, poisson_arrivals

    reqs = make_mix(n, seed, n2=n2, n3=n3, storage=storage, jacobi_cap=jacobi_cap)
    arr = poisson_arrivals(n, rate, seed + 1)
    acc_ids = accuracy_sample(reqs, every=accuracy_every)
    srv = Server(int(budget_gb * 1e9), round_sweeps=round_sweeps, keep_fields=acc_ids,
                 use_graph=use_graph)
    try:
        res, st = srv.serve(reqs, arr, time_limit=time_limit,
                            progress=_printer("serve: round {0}, {1} finished, {2} resident, "
                                              "{3:.0f} s") if progress else None)
    finally:
        srv.close()
    return {"workload": f"{n} Poisson arrivals at {rate:g}/s of the mix: 50% heat (200 steps) / "
                        f"50% Jacobi Laplace to 1e-6 (sweep cap {jacobi_cap}), 60% 2D N in "
                        f"{list(n2)}, 40% 3D N in {list(n3)}, {storage}, seed {seed}, HBM budget "
                        f"{budget_gb:g} GB",
            "rounds": st.rounds, "table_updates": st.table_updates,
            "table_update_s": round(st.table_update_s, 3), "max_resident": st.max_resident,
            "max_resident_GB": st.max_resident_bytes / 1e9,
            **outcome_stats(reqs, res, st.seconds),
            "statuses": {s_: sum(r.status == s_ for r in res) for s_ in ("done", "stagnated", "max_sweeps")},
            "accuracy_vs_fp64": accuracy_stats(reqs, res, progress)}
