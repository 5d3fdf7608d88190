"""CPU oracle of the organoid timestep -- test infrastructure only.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker / timed CPU baseline; the product library never uses it.
"""
