"""CPU oracle for the reconstruction path -- test infrastructure only.

Only the test-suite, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
baseline leg may import this package, and only as the checker.
"""
