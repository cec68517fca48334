"""MI355X (gfx950) state-vector engine behind the lightning.kokkos interfaces.

The compute path is the HIP library ``libplk_amd.so`` (C ABI in
``include/plk_amd.h``); ``lightning_kokkos_qubit_ops`` is the pybind11 module
that exposes the reference's class names over that C ABI.
"""
__version__ = "0.1.0"
