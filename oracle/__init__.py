"""FP64 CPU oracle of the phase-field hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check of __graft_entry__ and the CPU-baseline leg of
bench.py may import this package, and only as the checker of the HIP path.
"""
