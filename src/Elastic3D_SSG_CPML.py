"""``from src.Elastic3D_SSG_CPML import ElasticWAVE``, the import of the reference's
src/Elastic3D_SSG_CPML.py (3-D CPML: update_SSG), resolved to the MI355X implementation in
``seiswave_taichi_amd.Elastic3D_SSG_CPML``."""
from seiswave_taichi_amd.Elastic3D_SSG_CPML import ElasticWAVE

__all__ = ["ElasticWAVE"]
