"""Device-copy bandwidth ceiling, measured in the same process as the sweeps.

Counterpart of reference ``src/pyheat3d/roofline.py:8-29`` (torch ``copy_`` on the CUDA device):
here the copy is this library's own 16-byte streaming kernel (``ph3_copy_kernel``) and, as a
second figure, the runtime's device-to-device ``hipMemcpyAsync``. Read + write bytes are both
counted; the median of ``repeats`` is returned.
"""

from __future__ import annotations

import numpy as np

from . import _hip


def copy_bandwidth(nbytes: int = 1 << 31, iters: int = 10, repeats: int = 5,
                   method: str = "kernel") -> float:
    """Median copy bandwidth in GB/s (2 * nbytes per copy) of ``method`` in {kernel, memcpy}."""
    L = _hip.lib()
    src = _hip.malloc(nbytes)
    dst = _hip.malloc(nbytes)
    stream = _hip.Stream()
    ev0, ev1 = _hip.Event(), _hip.Event()
    try:
        def once() -> None:
            if method == "kernel":
                _hip.check(L.ph3_copy_kernel(src, dst, nbytes, stream.handle), "copy")
            else:
                _hip.check(L.ph3_memcpy_async(dst, src, nbytes, _hip.D2D, stream.handle), "copy")

        for _ in range(3):
            once()
        stream.sync()
        out = []
        for _ in range(repeats):
            ev0.record(stream.handle)
            for _ in range(iters):
                once()
            ev1.record(stream.handle)
            ev1.sync()
            ms = ev0.elapsed_ms(ev1)
            out.append(2.0 * nbytes * iters / (ms * 1e-3) / 1e9)
        return float(np.median(out))
    finally:
        ev0.close()
        ev1.close()
        stream.close()
        _hip.free(src)
        _hip.free(dst)
