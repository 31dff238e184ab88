"""``from src.Elastic2D_CPML import ElasticWAVE``, the import of the reference's
src/Elastic2D_CPML.py (2-D P-SV CPML: update_SSG_VS, update_SSG_SV, update_RSG, AEA), resolved to the MI355X implementation in
``seiswave_taichi_amd.Elastic2D_CPML``."""
from seiswave_taichi_amd.Elastic2D_CPML import ElasticWAVE

__all__ = ["ElasticWAVE"]
