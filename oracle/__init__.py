"""CPU oracle for the MPM substep -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package, and
only as the checker or as the timed CPU baseline. The product path (libmpm_hip.so and
anisotropicelastoplasticity_amd) never loads it.
"""
