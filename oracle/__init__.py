"""CPU FP64 oracle of Woo's DEM step — test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker (or the timed CPU
baseline), never as the thing measured or shipped. See ``oracle.h`` for the
reference functions it restates.
"""
from .oracle import Oracle, OrcParams, ENG, FIELD, build_oracle, liboracle_path, native_library  # noqa: F401
