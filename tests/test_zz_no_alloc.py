"""The residual entry point that allocates nothing, against the workspace one, through the C ABI.

Two calls in a row, the second on a smaller grid, must give the same FP64 bits as
``ph3_residual`` with a caller workspace (the same kernels and reduction order), and the value
the oracle computes to 1e-12.
"""

import ctypes

import numpy as np
import pytest

from oracle import numpy_ref as nr
from pyheat3d_amd import _hip

pytestmark = pytest.mark.gpu


def _upload(host: np.ndarray, storage: str, stream: _hip.Stream) -> int:
    n = host.size
    dev = _hip.malloc(n * _hip.STORAGE_BYTES[storage])
    if storage == "float32":
        _hip.memcpy(dev, host.ctypes.data, host.nbytes, _hip.H2D)
        return dev
    tmp = _hip.malloc(host.nbytes)
    try:
        _hip.memcpy(tmp, host.ctypes.data, host.nbytes, _hip.H2D)
        _hip.check(_hip.lib().ph3_encode(_hip.STORAGE_CODE[storage], tmp, dev, n, stream.handle),
                   "encode")
        stream.sync()
    finally:
        _hip.free(tmp)
    return dev


def test_noalloc_residual_equals_workspace_residual_on_two_calls():
    L = _hip.lib()
    assert L.ph3_residual_noalloc_scratch_bytes() >= L.ph3_residual_work_bytes()
    rng = np.random.default_rng(77)
    stream = _hip.Stream()
    out = _hip.malloc(16)
    work = _hip.malloc(int(L.ph3_residual_work_bytes()))
    cases = [((41, 50, 67), "dirichlet", "float32", True),    # first call: the larger grid
             ((9, 10, 11), "periodic", "bfloat16", False),    # second call: smaller
             ((33, 45), "dirichlet", "float16", True)]
    try:
        for shape, bc, storage, with_src in cases:
            print(f"[gpu] residual_noalloc {shape} {bc} {storage} source={with_src}", flush=True)
            d = len(shape)
            u = rng.random(shape, dtype=np.float32)
            src = (rng.random(shape, dtype=np.float32) * np.float32(1e-3)) if with_src else None
            du = _upload(u, storage, stream)
            db = _upload(src, "float32", stream) if with_src else 0
            nz, ny, nx = (1, *shape) if d == 2 else shape
            c0, c1 = 0.0, float(np.float32(1 / (2 * d)))
            args = (_hip.STORAGE_CODE[storage], int(bc == "periodic"), d, nz, ny, nx, c0, c1, du,
                    db or None)
            try:
                _hip.check(L.ph3_residual(*args, out, work, stream.handle), "residual")
                _hip.check(L.ph3_residual_noalloc(*args, out + 8, stream.handle),
                           "residual_noalloc")
                stream.sync()
                got = (ctypes.c_double * 2)()
                _hip.memcpy(ctypes.addressof(got), out, 16, _hip.D2H)
            finally:
                _hip.free(du)
                if db:
                    _hip.free(db)
            with_ws, no_alloc = got[0], got[1]
            assert np.float64(with_ws).tobytes() == np.float64(no_alloc).tobytes(), (shape,
                                                                                     with_ws,
                                                                                     no_alloc)
            want = nr.residual_fp64(nr.round_storage(u, storage), c1, bc, src)
            assert no_alloc == pytest.approx(want, rel=1e-12), shape
    finally:
        _hip.free(out)
        _hip.free(work)
        stream.close()
