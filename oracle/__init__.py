"""CPU FP64 oracle of the sweep engine - TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py use
this package, as the checker or the timed CPU baseline.
"""
