"""ctypes loader for the scalar C oracle ``oracle/_build/liboracle.so``. TEST INFRASTRUCTURE ONLY.

Parity status: ``ref_step_f32/f64`` pin :func:`oracle.numpy_ref.reference_step` (identical IEEE
op sequence); the codec is pinned by bit-pattern known answers in ``tests/test_oracle.py``.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_build" / "liboracle.so"
_lib = None

_F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build() -> Path:
    """Compile the C oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-f", str(_HERE / "Makefile")], check=True, cwd=_HERE)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        vp = ctypes.c_void_p
        ci, cd = ctypes.c_int, ctypes.c_double
        for name in ("ref_step_f32", "ref_step_f64"):
            f = getattr(L, name)
            f.argtypes = [vp, vp, ci, ci, ci, ci, cd, cd, vp, ci]
            f.restype = None
        L.ref_run_f32.argtypes = [vp, vp, vp, ci, ci, ci, ci, cd, cd, vp, ci, ci]
        L.ref_run_f32.restype = None
        L.ref_round_storage.argtypes = [vp, vp, vp, ctypes.c_size_t, ci]
        L.ref_round_storage.restype = None
        L.ref_residual_f64.argtypes = [vp, vp, ci, ci, ci, ci, cd, cd, ci]
        L.ref_residual_f64.restype = cd
        _lib = L
    return _lib


def _dims(u: np.ndarray) -> tuple[int, int, int, int]:
    if u.ndim == 2:
        return 2, 1, u.shape[0], u.shape[1]
    return 3, u.shape[0], u.shape[1], u.shape[2]


def _ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


def step(u: np.ndarray, c0: float, c1: float, bc: str, b: np.ndarray | None = None
         ) -> np.ndarray:
    """One update by the scalar C loop (float32 or float64 by ``u.dtype``)."""
    u = np.ascontiguousarray(u)
    o = np.empty_like(u)
    bb = None if b is None else np.ascontiguousarray(b, dtype=u.dtype)
    f = lib().ref_step_f64 if u.dtype == np.float64 else lib().ref_step_f32
    if u.dtype not in (np.float32, np.float64):
        raise TypeError(u.dtype)
    f(_ptr(u), _ptr(o), *_dims(u), c0, c1, _ptr(bb), int(bc == "periodic"))
    return o


def run_f32(u0: np.ndarray, c0: float, c1: float, bc: str, steps: int,
            b: np.ndarray | None = None) -> np.ndarray:
    u0 = np.ascontiguousarray(u0, dtype=np.float32)
    out = np.empty_like(u0)
    work = np.empty_like(u0)
    bb = None if b is None else np.ascontiguousarray(b, dtype=np.float32)
    lib().ref_run_f32(_ptr(u0), _ptr(out), _ptr(work), *_dims(u0), c0, c1, _ptr(bb),
                      int(bc == "periodic"), steps)
    return out


def round_storage(x: np.ndarray, storage: str) -> tuple[np.ndarray, np.ndarray]:
    """(bits uint16, widened float32) by the C codec."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    bits = np.empty(x.shape, np.uint16)
    back = np.empty(x.shape, np.float32)
    code = {"float16": 1, "bfloat16": 2}[storage]
    lib().ref_round_storage(_ptr(x), _ptr(bits), _ptr(back), x.size, code)
    return bits, back


def residual_fp64(u: np.ndarray, c1: float, bc: str, b: np.ndarray | None = None,
                  c0: float = 0.0) -> float:
    u = np.ascontiguousarray(u, dtype=np.float32)
    bb = None if b is None else np.ascontiguousarray(b, dtype=np.float32)
    return lib().ref_residual_f64(_ptr(u), _ptr(bb), *_dims(u), c0, c1, int(bc == "periodic"))
