"""qcdgpu_amd -- MI355X-native SU(N) Monte Carlo update and measurement path.

The compute path is libqcdhip.so (HIP kernels for gfx950 behind the C ABI in
include/qcd_hip.h); this package is its host-side mirror of QCDGPU's model
driver.  Importing it does not load the library; the first call does, and it
raises if the library is missing.
"""
from ._abi import QCDError, header_symbols  # noqa: F401
from .batch import Batch, polyakov_susceptibility  # noqa: F401
from .lattice import Lattice, device_info, prng_produce, read_qcg_header  # noqa: F401

__all__ = [
    "Batch", "Lattice", "QCDError", "device_info", "polyakov_susceptibility", "prng_produce", "header_symbols",
    "read_qcg_header",
]
