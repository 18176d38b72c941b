"""Benchmark of the stencil hot path on one MI355X: prints ONE JSON line on stdout.

Workload (the headline roofline configuration): 3D 7-point explicit FTCS heat, 1024^3 float32,
periodic boundaries, initial state = sum of three sine modes (wavenumbers 1, 2, 3 per axis, unit
amplitude) + 1e-3 * N(0,1) noise from seed 1, r = 0.15. One "step" is one sweep of the field
(one lattice update of all 1024^3 sites) by the headline variant ``hip-cdna4/zmarch``; the other
rungs (naive, tb<k>) are timed on the same field and reported beside it, each over at least
``--rung-steps`` steps (a tb4 run of 20 steps would be 5 launches, too few to time).

    python bench.py [--steps K] [--warmup W] [--n 1024] [--variant hip-cdna4/zmarch]

Timing: W untimed steps, then exactly K steps between two synchronisations; ``value`` is
lattice updates per second over that region (GLUPS, inputs resident in HBM). The dominant
kernel's average launch time comes from HIP events recorded on the stepper's own stream around
the same region. ``cpu_baseline`` times the NumPy oracle (one core) on a bounded sample.
``batch`` and ``serve`` carry the solves/s and p99 half of the metric on a bounded sample of the
mixed batch workload (``--batch-n``; see ``batch_sample``). ``config_128_dirichlet`` runs the
128^3 Dirichlet parity configuration in full on every float32 rung and on the NumPy oracle. ``comparators`` times the reference's
own torch-eager GPU rung on the same device and field (``torch_eager_comparator``).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
METRIC = ("solves/s + p99: batched server vs per-solve launches @1e-6; GLUPS, %HBM copy BW")
PROFILES = ROOT / "profiles"


def log(*a) -> None:
    print(*a, file=sys.stderr, flush=True)


def mode_amplitudes(u: np.ndarray, modes=(1, 2, 3)) -> list[float]:
    """Projection of a periodic n^3 field on prod_axes sin(2 pi m x_i), m in ``modes``."""
    n = u.shape[0]
    x = np.arange(n, dtype=np.float64) / n
    out = []
    for m in modes:
        s = np.sin(2 * np.pi * m * x)
        t = (u.reshape(-1, n) @ s.astype(np.float32)).astype(np.float64).reshape(n, n)
        out.append(float(s @ (t @ s)) / float(s @ s) ** 3)
    return out


def time_variant(key: str, u0: np.ndarray, stencil, steps: int, warmup: int,
                 amp0: list[float] | None = None) -> dict:
    from pyheat3d_amd import _hip
    from pyheat3d_amd.registry import VARIANTS, make_stepper

    var = VARIANTS[key]
    t_up = time.perf_counter()
    st = make_stepper(key, u0, stencil)  # allocation + host-to-device upload of u0
    t_up = time.perf_counter() - t_up
    tile = dict(st.tile)                  # the launch the library resolved (ph3_sweep_plan)
    try:
        k = var.steps_per_pass
        steps = max(k, steps - steps % k)
        st.step(max(warmup - warmup % k, k))
        st.synchronize()
        ev0, ev1 = _hip.Event(), _hip.Event()
        t0 = time.perf_counter()
        ev0.record(st.stream.handle)
        st.step(steps)
        ev1.record(st.stream.handle)
        st.synchronize()
        wall = time.perf_counter() - t0
        ev_ms = ev0.elapsed_ms(ev1)
        ev0.close()
        ev1.close()
        t_down = time.perf_counter()
        out = st.get()  # device-to-host download of the final field
        t_down = time.perf_counter() - t_down
        total_steps = st.steps_done
        finite = bool(np.isfinite(out).all())
        check = None
        if amp0 is not None:
            # every sine mode is an eigenvector of the update: amplitude ratio = g_m^steps with
            # g_m = c0 + 6 c1 cos(2 pi m / n) in the float32 coefficients the kernels use
            n = u0.shape[0]
            c0f, c1f = float(np.float32(stencil.c0)), float(np.float32(stencil.c1))
            amp = mode_amplitudes(out)
            errs = []
            for m, a0, a1 in zip((1, 2, 3), amp0, amp):
                g = c0f + 6.0 * c1f * math.cos(2 * math.pi * m / n)
                errs.append(abs(a1 / a0 / g ** total_steps - 1.0))
            check = {"steps": total_steps, "mode_amplitude_rel_err": errs}
        del out
    finally:
        st.close()
    cells = math.prod(u0.shape)
    bpu = 2.0 * var.bytes_per_value / k
    launches = steps // k
    return {"key": key, "steps": steps, "k": k, "wall_s": wall, "event_ms": ev_ms,
            "ms_per_step": 1e3 * wall / steps, "glups": cells * steps / wall / 1e9,
            "kernel_avg_ms": ev_ms / launches, "bytes_per_update": bpu,
            "algorithmic_bytes_per_launch": bpu * k * cells,
            "effective_GBs": cells * steps / wall * bpu / 1e9, "finite": finite,
            "check": check, "tile": tile,
            "host_transfer": {"upload_s": round(t_up, 4), "download_s": round(t_down, 4),
                              "bytes_each_way": int(u0.nbytes),
                              "glups_including_transfers": cells * steps / (wall + t_up + t_down) / 1e9}}


def load_traffic(workload: str) -> dict | None:
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of this workload."""
    p = PROFILES / f"pmc_{workload}.json"
    if not p.exists():
        return None
    d = json.loads(p.read_text())
    return d


def cpu_baseline(n: int, min_time: float) -> dict:
    """NumPy oracle (periodic slicing stepper, one core) on an n^3 sample of the workload."""
    from oracle import numpy_ref as nr
    from pyheat3d_amd.analytic import sine_modes_periodic
    from pyheat3d_amd.benchmark import cpu_model
    from pyheat3d_amd.problem import heat_stencil

    u = sine_modes_periodic(n, 3)
    st_ = heat_stencil(3, 0.15, "periodic")
    s = nr.NumpyStepper(u, st_.c0, st_.c1, "periodic")
    s.step(1)  # warm-up
    steps, t = 0, 0.0
    while t < min_time:
        t0 = time.perf_counter()
        s.step(1)
        t += time.perf_counter() - t0
        steps += 1
    return {"value": n ** 3 * steps / t / 1e9, "unit": "GLUPS", "cores": 1, "kind": "port",
            "sample": f"NumPy oracle periodic slicing stepper, {n}^3 fp32, {steps} steps "
                      f"({t:.1f} s) on one core of {cpu_model() or 'the host'} "
                      f"({os.cpu_count()} cores visible)", "cpu_model": cpu_model()}


def torch_eager_child(n: int, steps: int, warmup: int) -> int:
    """Child-process half of ``torch_eager_comparator``; never loads the HIP library.

    PyTorch-ROCm brings its own HIP runtime, and a second runtime in a process that already
    holds the device finds no GPU, so the comparator runs in a process of its own. Prints one
    JSON line: time per step and a SHA-256 of the final float32 field.
    """
    import hashlib

    import torch

    from pyheat3d_amd import analytic
    from pyheat3d_amd.problem import heat_stencil

    u0 = analytic.sine_modes_periodic(n, 3)
    stencil = heat_stencil(3, 0.15, "periodic")
    c0, c1 = float(stencil.c0), float(stencil.c1)

    def step(u, out):
        # backends/torch_backends.py:24-56: one roll pair per axis (x first), one expression
        s = (torch.roll(u, 1, 2) + torch.roll(u, -1, 2)) + (torch.roll(u, 1, 1) + torch.roll(u, -1, 1))
        s = s + (torch.roll(u, 1, 0) + torch.roll(u, -1, 0))
        out.copy_(c0 * u + c1 * s)

    dev = torch.device("cuda")
    a = torch.from_numpy(u0).to(dev)
    b = torch.empty_like(a)
    for _ in range(warmup):
        step(a, b)
        a, b = b, a
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        step(a, b)
        a, b = b, a
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    got = a.cpu().numpy()
    print(json.dumps({"ms_per_step": ms, "sha256": hashlib.sha256(got.tobytes()).hexdigest()}))
    return 0


def torch_eager_comparator(u0: np.ndarray, stencil, steps: int, warmup: int) -> dict:
    """The reference's GPU rung 1, ``torch-eager``, on this GPU (a same-device comparator).

    Restates the periodic path of ``backends/torch_backends.py:24-56`` (``_pairs`` with
    ``torch.roll`` per axis, x first, then the update written as one expression, each operator
    its own kernel) on PyTorch-ROCm, in a child process (``torch_eager_child``), timed with torch
    events over ``steps`` steps after ``warmup``. Its final field is compared bit for bit (by
    SHA-256) with the z-march stepper run for the same number of steps from the same field.
    (The reference's ``torch-compile`` rung generates Triton kernels and is not run here.)
    """
    import hashlib
    import subprocess

    from pyheat3d_amd.registry import make_stepper

    n = u0.shape[0]
    r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--torch-child",
                        str(n), str(steps), str(warmup)], capture_output=True, text=True,
                       timeout=600)
    if r.returncode != 0:
        return {"error": r.stderr.strip()[-400:]}
    child = json.loads(r.stdout.strip().splitlines()[-1])
    st = make_stepper("hip-cdna4/zmarch", u0, stencil)
    try:
        st.step(warmup + steps)
        ours = st.get()
    finally:
        st.close()
    cells = math.prod(u0.shape)
    ms = child["ms_per_step"]
    return {"what": "reference torch-eager rung (torch.roll per axis, one expression) on "
                    "PyTorch-ROCm, same GPU and field, child process",
            "glups": round(cells / ms / 1e6, 2), "ms_per_step": round(ms, 3), "steps": steps,
            "bitwise_equal_to_zmarch": hashlib.sha256(ours.tobytes()).hexdigest() == child["sha256"]}


def _cpu_solve_seconds(args: tuple) -> float:
    """One-core NumPy-oracle time of one batch request (module level, for a process pool).

    A solve is timed in full when it is at most 5e7 site updates; otherwise 5 steps (heat) or
    20 sweeps plus one FP64 residual (Jacobi) are timed and scaled to the step count of the
    GPU run (Jacobi: with a residual every 64 sweeps).
    """
    from oracle import numpy_ref as nr
    from pyheat3d_amd.analytic import splitmix_field

    shape, seed, kind, c0, c1, mode, sweeps = args
    s = nr.NumpyStepper(splitmix_field(shape, seed, kind), c0, c1, "dirichlet")
    sites = math.prod(shape)
    if mode == "heat":
        part = sweeps if sites * sweeps <= 5e7 else 5
        t0 = time.perf_counter()
        s.step(part)
        return (time.perf_counter() - t0) * sweeps / part
    part = min(20, sweeps) if sites * sweeps > 5e7 else sweeps
    t0 = time.perf_counter()
    s.step(part)
    t_sweep = (time.perf_counter() - t0) / max(part, 1)
    t0 = time.perf_counter()
    nr.residual_fp64(s.get(), c1, "dirichlet")
    t_res = time.perf_counter() - t0
    return t_sweep * sweeps + t_res * (sweeps // 64 + 1)


def cpu_batch_baseline(reqs: list, sweeps: dict[int, int]) -> dict:
    """The NumPy oracle on the per-solve sample of the batch mix: one core, then every core."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor

    jobs = []
    for q in reqs:
        st = q.stencil()
        jobs.append((q.shape, q.seed, q.kind, st.c0, st.c1, q.mode, sweeps[q.rid]))
    t0 = time.perf_counter()
    one = [_cpu_solve_seconds(j) for j in jobs]
    measured = time.perf_counter() - t0
    cores = max(1, min(16, os.cpu_count() or 1))
    # spawned workers: fresh interpreters that never touch the GPU this process holds
    with ProcessPoolExecutor(cores, mp_context=multiprocessing.get_context("spawn")) as ex:
        many = list(ex.map(_cpu_solve_seconds, jobs))
    return {"kind": "port", "solves": len(jobs),
            "solves_per_s_1core": round(len(jobs) / sum(one), 4),
            "solves_per_s_all_cores": round(len(jobs) * cores / sum(many), 4),
            "cores": cores,
            "sample": "NumPy oracle (Dirichlet slicing stepper) on the per-solve sample, the "
                      "GPU run's step counts; solves over 5e7 site updates timed on 5 steps "
                      "(heat) or 20 sweeps + one FP64 residual (Jacobi) and scaled; all-core "
                      f"figure assumes independent solves on {cores} processes ({measured:.1f} s "
                      "measured)"}


#: the bench's sample of the mixed batch workload: config 4's size ranges, a sweep cap that
#: keeps the default run within minutes
BATCH_N2, BATCH_N3, BATCH_CAP = (64, 2048), (32, 256), 16384


def batch_sample(n: int) -> tuple[dict, dict]:
    """The solves/s + p99 half of the metric on a bounded sample of the batch mix.

    ``n`` mixed solves drawn from the mixed workload's full size ranges (60% 2D N in
    [64, 2048], 40% 3D N in [32, 256], half 200 heat steps, half Jacobi Laplace to a 1e-6
    relative FP64 residual with a sweep cap of ``BATCH_CAP``, float32): all submitted at once,
    batched graph rounds vs one launch per pass per solve on the batch's own fused-step rung
    (every 16th solve; also with the stepper's replayed graphs), then the same mix served at
    1000 Poisson arrivals per second. Solves that reached their target are counted apart from
    those that stagnated at the float32 floor or hit the cap; every 32nd result of bounded size
    is compared with its FP64 NumPy solution. The full mixes run under
    ``python -m pyheat3d_amd batch|serve``.
    """
    from pyheat3d_amd.batch import make_mix
    from pyheat3d_amd.benchmark import batch_benchmark, serve_benchmark

    t0 = time.perf_counter()
    b = batch_benchmark(n, seed=2024, n2=BATCH_N2, n3=BATCH_N3, frac_jacobi=0.5,
                        baseline_every=16, jacobi_cap=BATCH_CAP, accuracy_every=32)
    bb, ps = b["batched"], b["per_solve"]
    log(f"batch: {bb['solves_per_s']:.1f} solves/s batched ({bb['reached_target']['solves']} "
        f"reached the target) vs {ps['solves_per_s']:.2f} per-solve ({b['speedup']:.1f}x; "
        f"{b['per_solve_graphs']['solves_per_s']:.2f} with replayed graphs, "
        f"{b['speedup_vs_graphs']:.1f}x) in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    sample = make_mix(n, 2024, n2=BATCH_N2, n3=BATCH_N3, frac_jacobi=0.5,
                      jacobi_cap=BATCH_CAP)[::16]
    cpu = cpu_batch_baseline(sample, b["sample_sweeps"])
    log(f"batch cpu baseline: {cpu['solves_per_s_1core']:.3f} solves/s on 1 core, "
        f"{cpu['solves_per_s_all_cores']:.3f} on {cpu['cores']} in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    s = serve_benchmark(n, rate=1000.0, seed=2024, n2=BATCH_N2, n3=BATCH_N3,
                        jacobi_cap=BATCH_CAP, accuracy_every=32)
    log(f"serve: {s['solves_per_s']:.1f} solves/s, p50 {s['p50_ms']:.0f} ms, "
        f"p99 {s['p99_ms']:.0f} ms in {time.perf_counter() - t0:.1f} s")

    def slim(o: dict) -> dict:
        keep = ("solves", "solves_per_s", "p50_ms", "p99_ms", "max_ms")
        return {k: (round(v, 3) if isinstance(v, float) else v) for k, v in o.items() if k in keep}

    batch = {"workload": b["workload"],
             "batched": {**slim(bb), "effective_glups": round(bb["effective_glups"], 1),
                         **{g: slim(bb[g]) for g in ("reached_target", "stagnated", "max_sweeps")
                            if g in bb}},
             "per_solve": {**slim(ps), "sample": ps["solves"], "rung": ps["rung"],
                           **{g: slim(ps[g]) for g in ("reached_target", "stagnated", "max_sweeps")
                              if g in ps}},
             "speedup": round(b["speedup"], 2),
             "speedup_reached_target": round(b.get("speedup_reached_target", float("nan")), 2),
             "per_solve_graphs_solves_per_s": round(b["per_solve_graphs"]["solves_per_s"], 3),
             "speedup_vs_graphs": round(b["speedup_vs_graphs"], 2),
             "jacobi": bb["jacobi"],
             "baseline_agrees_on_steps_and_status": b["baseline_agrees_on_steps_and_status"],
             "accuracy_vs_fp64": b["accuracy_vs_fp64"],
             "cpu_baseline": cpu}
    serve = {"workload": s["workload"], **slim(s),
             **{g: slim(s[g]) for g in ("reached_target", "stagnated", "max_sweeps") if g in s},
             "statuses": s["statuses"], "accuracy_vs_fp64": s["accuracy_vs_fp64"]}
    return batch, serve


def config1_leg(keys: list[str], steps: int = 100, repeats: int = 5) -> dict:
    """The parity configuration at its full size: 128^3 float32, Dirichlet walls 0, interior
    U[0,1) from default_rng(0), r = 0.15, 100 steps. The NumPy oracle runs it in full on one core
    (the CPU baseline of this configuration and the bitwise check of every rung); each GPU rung
    is timed as the median of ``repeats`` runs of ``steps`` steps (interior sites)."""
    from oracle import numpy_ref as nr
    from pyheat3d_amd.analytic import random_dirichlet
    from pyheat3d_amd.problem import heat_stencil
    from pyheat3d_amd.registry import make_stepper

    u0 = random_dirichlet((128, 128, 128), 0)
    st = heat_stencil(3, 0.15)
    interior = 126 ** 3
    t0 = time.perf_counter()
    ref = nr.emulated_reference(u0, st.c0, st.c1, "dirichlet", steps)
    cpu_s = time.perf_counter() - t0
    out = {"workload": f"3D 7-point FTCS heat 128^3 fp32 Dirichlet (walls 0, interior U[0,1) "
                       f"seed 0), r=0.15, {steps} steps",
           "cpu_baseline": {"value": round(interior * steps / cpu_s / 1e9, 4), "unit": "GLUPS",
                            "cores": 1, "kind": "port",
                            "sample": f"NumPy oracle, the whole run ({cpu_s:.2f} s)"},
           "rungs": {}}
    for key in keys:
        s = make_stepper(key, u0, st)
        try:
            s.step(steps)
            s.synchronize()
            bitwise = bool(np.array_equal(s.get(), ref))
            times = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                s.step(steps)
                s.synchronize()
                times.append(time.perf_counter() - t0)
        finally:
            s.close()
        out["rungs"][key] = {"glups": round(interior * steps / float(np.median(times)) / 1e9, 2),
                             "bitwise_equal_to_oracle": bitwise}
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--variant", default="hip-cdna4/zmarch")
    ap.add_argument("--others", default="hip-cdna4/naive,hip-cdna4/tb2,hip-cdna4/tb3,hip-cdna4/tb4",
                    help="comma-separated rungs timed beside the headline ('' for none)")
    ap.add_argument("--rung-steps", type=int, default=240,
                    help="the other rungs run max(--steps, this) steps after max(--warmup, 24), "
                         "so that a rung of k steps per launch is timed over enough launches")
    ap.add_argument("--cpu-n", type=int, default=512)
    ap.add_argument("--cpu-min-time", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-copy", action="store_true", help="skip the copy-bandwidth probe")
    ap.add_argument("--no-torch", action="store_true", help="skip the torch-eager comparator")
    ap.add_argument("--torch-child", nargs=3, type=int, metavar=("N", "STEPS", "WARMUP"),
                    help=argparse.SUPPRESS)
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the 128^3 Dirichlet parity configuration")
    ap.add_argument("--batch-n", type=int, default=512,
                    help="solves in the batched-vs-per-solve and served sample (0: skip)")
    a = ap.parse_args()
    if a.torch_child:
        return torch_eager_child(*a.torch_child)
    if a.gpus != 1:
        print(f"bench.py: this benchmark runs on one GPU only; --gpus {a.gpus} is not supported "
              "(the multi-GPU z-slab design is documented, not built)", file=sys.stderr)
        return 2

    from pyheat3d_amd import _hip, analytic, roofline
    from pyheat3d_amd.problem import heat_stencil
    from pyheat3d_amd.registry import availability
    from pyheat3d_amd.traffic import SPEC_BW_GBS

    ok, why = availability(a.variant)
    if not ok:
        raise SystemExit(f"{a.variant} unavailable: {why}")
    info = _hip.device_info(0)
    log(f"device: {info['name']} {info['arch']} CUs={info['cu_count']}")

    copy_kernel = copy_memcpy = None
    if not a.no_copy:
        copy_kernel = roofline.copy_bandwidth(1 << 31, method="kernel")
        copy_memcpy = roofline.copy_bandwidth(1 << 31, method="memcpy")
        log(f"copy bandwidth: kernel {copy_kernel:.0f} GB/s, hipMemcpy D2D {copy_memcpy:.0f} GB/s")

    t0 = time.perf_counter()
    u0 = analytic.sine_modes_periodic(a.n, 3)
    log(f"init {a.n}^3 in {time.perf_counter() - t0:.1f} s")
    stencil = heat_stencil(3, 0.15, "periodic")
    amp0 = mode_amplitudes(u0)

    head = time_variant(a.variant, u0, stencil, a.steps, a.warmup, amp0)
    log(f"{a.variant}: {head['glups']:.2f} GLUPS, {head['effective_GBs']:.0f} GB/s, "
        f"kernel {head['kernel_avg_ms']:.3f} ms")
    others = {}
    for key in [x for x in a.others.split(",") if x]:
        okk, whyk = availability(key)
        if not okk:
            others[key] = {"unavailable": whyk}
            continue
        r = time_variant(key, u0, stencil, max(a.steps, a.rung_steps), max(a.warmup, 24), amp0)
        log(f"{key}: {r['glups']:.2f} GLUPS ({r['effective_GBs']:.0f} GB/s minimal model)")
        others[key] = {"glups": round(r["glups"], 3), "ms_per_step": round(r["ms_per_step"], 4),
                       "minimal_GBs": round(r["effective_GBs"], 1), "k": r["k"],
                       "single_step_equiv_GBs": round(r["glups"] * 8, 1),
                       "tile": r["tile"],
                       "max_mode_amplitude_rel_err": max(r["check"]["mode_amplitude_rel_err"])}
        if r["k"] > 1:   # the halo over-fetch model of the reference (traffic.py:38-50)
            from pyheat3d_amd.traffic import halo_model_bytes

            hb = halo_model_bytes(3, r["k"], 4, r["tile"])
            others[key]["halo_model_bytes_per_unit"] = round(hb, 4)
            others[key]["halo_model_GBs"] = round(r["glups"] * hb, 1)
        pmc = load_traffic(f"{key.split('/')[1]}_{a.n}")
        if pmc and pmc.get("variant") == key and pmc.get("n") == a.n:
            # HBM bytes per launch from the committed counter runs, as a rate at this run's
            # kernel time, and per site update
            t = pmc["hbm_bytes_per_launch"]
            others[key]["counter_traffic"] = {
                "bytes_per_launch": round(t), "over_algorithmic": round(pmc["traffic_over_algorithmic"], 4),
                "bytes_per_update": round(t / (a.n ** 3 * r["k"]), 4),
                "GBs_at_this_run": round(t / (r["kernel_avg_ms"] * 1e-3) / 1e9, 1)}
    torch_cmp = None
    if not a.no_torch:
        torch_cmp = torch_eager_comparator(u0, stencil, 20, 3)
        if "error" in torch_cmp:
            log(f"torch-eager comparator failed: {torch_cmp['error']}")
        else:
            log(f"torch-eager comparator: {torch_cmp['glups']:.1f} GLUPS, bitwise equal to "
                f"zmarch: {torch_cmp['bitwise_equal_to_zmarch']}")
    del u0

    achieved = head["algorithmic_bytes_per_launch"] / (head["kernel_avg_ms"] * 1e-3) / 1e9
    traffic = None
    pmc = load_traffic(f"zmarch_{a.n}")
    if pmc and pmc.get("variant") == a.variant and pmc.get("n") == a.n:
        traffic = round(pmc["hbm_bytes_per_launch"])
    config1 = None
    if not a.no_config1:
        config1 = config1_leg(["hip-cdna4/naive", "hip-cdna4/zmarch", "hip-cdna4/tb2",
                               "hip-cdna4/tb3", "hip-cdna4/tb4"])
        log("config 128^3 Dirichlet 100 steps: " + ", ".join(
            f"{k.split('/')[1]} {v['glups']:.0f}{'' if v['bitwise_equal_to_oracle'] else ' MISMATCH'}"
            for k, v in config1["rungs"].items()) + f" GLUPS; NumPy oracle "
            f"{config1['cpu_baseline']['value']:.3f} GLUPS on one core")
    batch = serve = None
    if a.batch_n > 0:
        batch, serve = batch_sample(a.batch_n)
    cpu = None if a.no_cpu_baseline else cpu_baseline(a.cpu_n, a.cpu_min_time)

    line = {
        "metric": METRIC, "value": round(head["glups"], 3), "unit": "GLUPS", "n_gpus": 1,
        "steps": head["steps"], "warmup": a.warmup, "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (3 sine modes + 1e-3 N(0,1) noise, seed 1)",
        "config": {"workload": f"3D 7-point FTCS heat {a.n}^3 fp32 periodic r=0.15, "
                               f"variant {a.variant}", "grid": [a.n] * 3, "r": 0.15,
                   "variant": a.variant, "parallelism": "single GPU"},
        "pct_copy_bw": None if not copy_kernel else round(100 * head["effective_GBs"] / copy_kernel, 1),
        "copy_bw_GBs": {"kernel": copy_kernel and round(copy_kernel, 1),
                        "hipMemcpyDtoD": copy_memcpy and round(copy_memcpy, 1)},
        "pct_spec_8TBs": round(100 * head["effective_GBs"] / SPEC_BW_GBS, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": SPEC_BW_GBS,
                     "unit": "GB/s", "frac": round(achieved / SPEC_BW_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
        "check": {"finite": head["finite"], **head["check"],
                  "what": "amplitude of each sine mode after the run vs the closed form g_m^steps"},
        "host_transfer": head["host_transfer"],
        "rungs": others,
        "config_128_dirichlet": config1,
        "comparators": {"torch-eager": torch_cmp},
        "batch": batch,
        "serve": serve,
        "device": {"name": info["name"], "arch": info["arch"], "cu_count": info["cu_count"]},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
