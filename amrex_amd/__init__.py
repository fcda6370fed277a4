"""MI355X-native MLMG (AMReX cell-centred multigrid) for gfx950.

The solver is C++/HIP (``csrc/``), reached through the C ABI in
``include/amrex_mlmg.h``; ``mlmg`` mirrors the reference call surface in
Python, ``workload`` generates the seeded benchmark inputs.
"""
__all__ = ["capi", "mlmg", "workload"]
