"""MI355X (gfx950) single-GPU spiking-network simulator with the reference's network API.

The simulation runs in hand-written HIP kernels behind a C ABI (include/spice_hip.h,
spice_amd/libspice_hip.so); this package is the Python host mirror of that API.
"""
from .models import MODELS, PLASTIC
from .serve import Server
from .snn import Batch, Layout, Snn

__all__ = ["Batch", "Layout", "Server", "Snn", "MODELS", "PLASTIC"]
