"""FP64 CPU oracle for the beamz FDTD hot path.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or, for ``cpu_baseline``, as the timed stand-in of the reference's own
NumPy path). The product package ``beamz_amd`` never imports it.

Every function restates the reference algorithm with the same array-op order,
citing the reference file:line it follows (paths relative to the reference
repository root). How each function is pinned (golden KATs, closed forms,
invariants, brute force) is listed in the header of each module.
"""
