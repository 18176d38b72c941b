"""Benchmark entry points mirroring the reference's ``bench/run.py`` and ``bench/tune.py``."""
