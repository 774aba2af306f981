"""TEST INFRASTRUCTURE ONLY -- CPU oracles of the image structure function.

Imported by tests/, the smoke check and bench.py's cpu_baseline leg, never by
the fastddm_amd package. ``oracle_np`` is the FP64 NumPy oracle; ``port`` loads
``liboracle.so``, the FP64 C restatement of the reference's C++ core.
"""
