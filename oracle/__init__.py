"""CPU oracle for the uplift-erosion loop (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this package, and only as the checker.  The product (fastlem_amd) never
imports it.
"""
