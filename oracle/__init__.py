"""CPU oracle of the tracking hot path -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the parity checker.  The product package never imports it.
"""
