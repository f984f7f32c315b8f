"""CPU oracle for the VOSS hot path -- test infrastructure only.

Imported only by tests/, by ``__graft_entry__.smoke()`` (to compare its
result with) and by ``bench.py``'s cpu_baseline leg.  The product never
imports it.
"""
