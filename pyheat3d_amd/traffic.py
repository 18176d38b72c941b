"""Traffic and FLOP models that turn lattice updates per second into bandwidth and roofline %.

Minimal model (reference ``src/pyheat3d/traffic.py:1-35``): a sweep advancing the field by ``k``
steps must read every value once and write it once, so one site update costs
``2 * bytes_per_value / k`` bytes (8 B fp32, 4 B fp16/bf16, 8/k B for ``tb<k>``). A source term
adds one read of ``bytes_source / k``. Halo re-reads are not counted, so the model is a lower
bound on real traffic; :func:`tb_model_bytes` gives the overlapped-tiling halo estimate.

Ceilings are MI355X figures: the 8 TB/s HBM3E spec, and a device-copy bandwidth that the
benchmark measures in the same process (``pyheat3d_amd.roofline.copy_bandwidth``).
"""

from __future__ import annotations

#: MI355X HBM3E peak (8 stacks, 8 TB/s), in GB/s
SPEC_BW_GBS = 8000.0
#: MI355X device-copy bandwidth, GB/s: this library's copy kernel on 2 GiB (6.12-6.15 TB/s
#: measured over several boxes). The stand-in when no measurement is at hand, like the
#: reference's measured-once ``COPY_BW_GBS`` for its GPU (``traffic.py``).
COPY_BW_GBS = 6140.0
#: MI355X dense FP32 vector peak, TFLOP/s (the stencil never gets near it: ~1 FLOP/B)
SPEC_FP32_TFLOPS = 157.3
#: bytes per stored value by storage name
BYTES = {"float32": 4, "float16": 2, "bfloat16": 2, "float64": 8}


def minimal_bytes(bytes_per_value: int, steps_per_pass: int = 1, source_bytes: int = 0) -> float:
    """Minimal DRAM bytes per site update: one read and one write per pass (+ source read)."""
    return (2.0 * bytes_per_value + source_bytes) / steps_per_pass


def flops_per_update(dims: int) -> int:
    """(2d-1) neighbour adds, then c1*S, c0*u and their sum."""
    return 2 * dims + 2


def roofline_mlups(bw_gbs: float, bytes_per_update: float) -> float:
    """Upper bound in million lattice updates per second for a bandwidth ceiling."""
    return bw_gbs * 1e3 / bytes_per_update


def tb_model_bytes(dims: int, K: int, bytes_per_value: int, tx: int, ty: int, zch: int) -> float:
    """Bytes per update of overlapped temporal blocking with no cache reuse of halos.

    Each tile of ``tx`` (x) by ``ty`` (y, 3D only) reads its extended tile, including a ``K``
    deep halo on every tiled side and ``K`` extra planes at both ends of its ``zch`` march, and
    writes the inner ``(tx-2K)(ty-2K)`` part once per ``K`` steps.
    """
    if dims == 3:
        over = tx * ty / ((tx - 2.0 * K) * (ty - 2.0 * K))
    else:
        over = tx / (tx - 2.0 * K)
    over *= (zch + 2.0 * K) / zch
    return bytes_per_value * (over + 1.0) / K


def halo_model_bytes(dims: int, K: int, bytes_per_value: int, plan: dict) -> float:
    """:func:`tb_model_bytes` for the tile a temporal-blocking launch actually uses (its
    ``ph3_plan``): the extended tile ``tile_x`` x ``tile_y`` of which ``out_x`` x ``out_y`` is
    written (the x halo is rounded up to whole lane vectors), and ``K`` extra planes (3D) or rows
    (2D) at both ends of each ``zch`` march."""
    over = plan["tile_x"] / plan["out_x"]
    if dims == 3:
        over *= plan["tile_y"] / plan["out_y"]
    zch = plan["zch"]
    over *= (zch + 2.0 * K) / zch
    return bytes_per_value * (over + 1.0) / K
