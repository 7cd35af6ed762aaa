"""CPU oracle for the patch-search path (test infrastructure only).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker. The product never imports it.
"""
