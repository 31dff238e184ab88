"""``from src.Elastic3D_SSG_Tensor import ElasticWAVE``, the import of the reference's
src/Elastic3D_SSG_Tensor.py (3-D CPML, moment tensor, receiver array: update_SSG), resolved to the MI355X implementation in
``seiswave_taichi_amd.Elastic3D_SSG_Tensor``."""
from seiswave_taichi_amd.Elastic3D_SSG_Tensor import ElasticWAVE

__all__ = ["ElasticWAVE"]
