"""CPU oracle for the pyheat3d stencil hot path. TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import or run anything in this package, and only as the checker or the timed CPU baseline. The
``pyheat3d_amd`` product path never imports it and has no CPU fallback.

* :mod:`oracle.numpy_ref` restates the reference's NumPy ``slicing`` backend
  (``src/pyheat3d/backends/numpy_ref.py``), its storage-rounding reference and its FP64
  validation studies (``src/pyheat3d/validation.py``).
* ``oracle/stencil_ref.c`` is a scalar per-site C restatement of the same update, compiled by
  ``oracle/Makefile`` into ``oracle/_build/liboracle.so``; :mod:`oracle.cref` loads it.

Parity status of each function is listed in the header of the module that defines it.
"""
