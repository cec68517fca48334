"""MI355X (gfx950) state-vector engine behind the lightning.kokkos interfaces.

The compute path is the HIP library ``libplk_amd.so`` (C ABI in
``include/plk_amd.h``); ``lightning_kokkos_qubit_ops`` is the pybind11 module
that exposes the reference's class names over that C ABI.
"""
__version__ = "0.1.0"


def __getattr__(name):
    # The plugin entry point "lightning.kokkos" names LightningKokkos; the
    # device module (and with it the HIP library) loads on first use.
    if name in ("LightningKokkos", "LightningKokkosAMD"):
        from .device import LightningKokkosAMD

        return LightningKokkosAMD
    raise AttributeError(name)
