"""CPU oracle for the MLMG hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It restates the reference algorithm on
the CPU (``mlmg_oracle.cpp``) and is the checker, never the product path.
"""
