"""``from src.Elastic3D_SSG_SPML import ElasticWAVE``, the import of the reference's
src/Elastic3D_SSG_SPML.py (3-D split-field PML: update_SSG_SPML), resolved to the MI355X implementation in
``seiswave_taichi_amd.Elastic3D_SSG_SPML``."""
from seiswave_taichi_amd.Elastic3D_SSG_SPML import ElasticWAVE

__all__ = ["ElasticWAVE"]
