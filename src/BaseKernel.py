"""``from src.BaseKernel import Ricker, Ricker2`` (reference src/BaseKernel.py:4-19,
imported by Relay_wave.py:5), resolved to the host wavelets of
``seiswave_taichi_amd.BaseKernel``."""
from seiswave_taichi_amd.BaseKernel import Ricker, Ricker2

__all__ = ["Ricker", "Ricker2"]
