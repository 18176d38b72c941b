"""Variant registry with the ``hip-cdna4/*`` family (reference ``src/pyheat3d/registry.py:1-209``).

Keys are ``"<backend>/<variant>"``; factories import lazily, so listing the registry needs no
GPU. ``make_stepper`` is the single entry point every caller (CLI, benchmark, validation, tests)
uses. The reference's libcuda probe (``registry.py:164-176``) is replaced by a HIP probe that
requires a gfx950 device.
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass, field
from functools import cache, lru_cache

import numpy as np

from .backends.base import Stepper
from .problem import Stencil

BACKENDS = ("hip-cdna4",)
COLORS = {"hip-cdna4": "#ed1c24"}
#: device architectures the library is compiled for
ARCHS = ("gfx950",)


@dataclass(frozen=True)
class Variant:
    """One rung of the ladder (reference ``registry.py:33-55``)."""

    backend: str
    name: str
    factory: Callable[[], type[Stepper]] = field(repr=False, compare=False)
    dims: tuple[int, ...] = (2, 3)
    storage: str = "float32"
    steps_per_pass: int = 1
    gpu: bool = True
    source: bool = False
    rung: int = 0
    description: str = ""

    @property
    def key(self) -> str:
        return f"{self.backend}/{self.name}"

    @property
    def bytes_per_value(self) -> int:
        return {"float32": 4, "float16": 2, "bfloat16": 2, "float64": 8}.get(self.storage, 4)


def _hip(kind: str, storage: str = "float32", K: int = 1):
    def f():
        from .backends import hip_backend

        return hip_backend.make(kind, storage, K)

    return f


_SHORT = {"float32": "", "float16": "-fp16", "bfloat16": "-bf16"}
_ST = ("float32", "float16", "bfloat16")


def _build() -> dict[str, Variant]:
    v: list[Variant] = []
    for st in _ST:
        v.append(Variant("hip-cdna4", "naive" + _SHORT[st], _hip("naive", st), storage=st,
                         source=True, rung=3,
                         description=f"HIP wave64, one site per lane, {st} storage"))
    for st in _ST:
        v.append(Variant("hip-cdna4", "zmarch" + _SHORT[st], _hip("zmarch", st), dims=(3,),
                         storage=st, source=True, rung=4,
                         description=f"HIP 2.5D z-march: register plane queue, 128-bit rows, "
                                     f"DPP x-halo, {st} storage"))
    for st in _ST:
        v.append(Variant("hip-cdna4", "ymarch" + _SHORT[st], _hip("tb", st, 1), dims=(2,),
                         storage=st, source=True, rung=4,
                         description=f"HIP 2D y-march with a register row window, {st} storage"))
    for K in (2, 3, 4):
        v.append(Variant("hip-cdna4", f"tb{K}", _hip("tb", "float32", K), steps_per_pass=K,
                         source=True, rung=5,
                         description=f"HIP temporal blocking, {K} steps per HBM pass"))
    v.append(Variant("hip-cdna4", "tb8", _hip("tb", "float32", 8), dims=(2,), steps_per_pass=8,
                     source=True, rung=5, description="HIP temporal blocking, 8 steps per pass"))
    for st in ("float16", "bfloat16"):
        v.append(Variant("hip-cdna4", "tb4" + _SHORT[st], _hip("tb", st, 4), storage=st,
                         steps_per_pass=4, source=True, rung=6,
                         description=f"HIP temporal blocking (k=4), {st} storage"))
    return {x.key: x for x in v}


VARIANTS: dict[str, Variant] = _build()


@lru_cache(maxsize=1)
def hip_available() -> bool:
    """True if libpyheat3d_hip.so loads and device 0 is a gfx950 (MI355X-class) GPU."""
    return hip_unavailable_reason() == ""


@lru_cache(maxsize=1)
def hip_unavailable_reason() -> str:
    try:
        from . import _hip

        if _hip.device_count() < 1:
            return "no HIP GPU"
        arch = str(_hip.device_info(0)["arch"]).split(":")[0]
        if arch not in ARCHS:
            return f"device arch {arch} is not one of {ARCHS}"
    except Exception as e:  # library missing or broken
        return f"{type(e).__name__}: {e}".splitlines()[0][:160]
    return ""


@cache
def availability(key: str) -> tuple[bool, str]:
    """``(ok, reason)`` for one variant (reference ``registry.py:179-189``)."""
    var = VARIANTS[key]
    if var.gpu and not hip_available():
        return False, hip_unavailable_reason()
    try:
        var.factory()
    except Exception as e:
        return False, f"{type(e).__name__}: {e}".splitlines()[0][:160]
    return True, ""


def available_variants(dims: int | None = None) -> list[str]:
    return [k for k, v in VARIANTS.items()
            if (dims is None or dims in v.dims) and availability(k)[0]]


def make_stepper(key: str, u0: np.ndarray, stencil: Stencil,
                 source: np.ndarray | None = None) -> Stepper:
    """Instantiate a variant on an initial state (reference ``registry.py:198-209``)."""
    var = VARIANTS[key]
    if u0.ndim not in var.dims:
        raise ValueError(f"{key} supports dims {var.dims}, got {u0.ndim}")
    if source is not None and not var.source:
        raise ValueError(f"{key} does not support a source term")
    cls = var.factory()
    st = cls(u0, stencil, source)
    st.steps_per_pass = var.steps_per_pass
    return st
