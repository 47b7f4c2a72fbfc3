"""CPU oracle for the QC-LDPC codec -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg,
to check results against.  The product package never imports it.
"""
