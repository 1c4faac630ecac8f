"""CPU oracle for the Hotspice stepping path.

TEST INFRASTRUCTURE ONLY. This package restates the reference algorithm in FP64 NumPy
(model.py, lattices.py) with replayable Philox4x32-10 streams (philox.py). Only tests/,
__graft_entry__.smoke() and the cpu_baseline leg of bench.py may import it, and only as
the checker. The product package ``hotspice_amd`` never imports, links or runs it.
"""
