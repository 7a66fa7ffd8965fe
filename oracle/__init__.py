"""CPU oracle for the free-wake induced-velocity path (test infrastructure only).

Loaded only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg.
"""
