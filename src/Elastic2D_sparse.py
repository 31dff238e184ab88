"""``from src.Elastic2D_sparse import ElasticWAVE``, the import of the reference's
src/Elastic2D_sparse.py (2-D CPML, sparse memory variables), resolved to the MI355X implementation in
``seiswave_taichi_amd.Elastic2D_sparse``."""
from seiswave_taichi_amd.Elastic2D_sparse import ElasticWAVE

__all__ = ["ElasticWAVE"]
