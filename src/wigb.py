"""``from src.wigb import *`` (reference src/wigb.py, imported by
test3d_homegenous_tensor.py:5), resolved to ``seiswave_taichi_amd.wigb``."""
from seiswave_taichi_amd.wigb import wigb

__all__ = ["wigb"]
