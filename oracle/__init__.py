"""CPU oracle of the CRTomo forward / Jacobian / CG path.

TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import it.
"""
