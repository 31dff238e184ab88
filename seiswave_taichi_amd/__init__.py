"""seiswave_taichi_amd -- MI355X-native elastic staggered-grid finite differences.

Drop-in replacements for the ElasticWAVE classes of SeisWave_Taichi's ``src``
package, with the time step in hand-written HIP kernels for gfx950
(libseiswave.so, C ABI in include/seiswave.h):

    from seiswave_taichi_amd.Elastic3D_SSG_CPML import ElasticWAVE
    from seiswave_taichi_amd.Elastic3D_SSG_Tensor import ElasticWAVE
    from seiswave_taichi_amd.Elastic3D_SSG_SPML import ElasticWAVE
    from seiswave_taichi_amd.Elastic2D_CPML import ElasticWAVE
    from seiswave_taichi_amd.Elastic2D_sparse import ElasticWAVE

``install_src_alias()`` makes ``from src.<Module> import ElasticWAVE`` resolve
to these modules, so reference driver scripts run unchanged.
"""
import importlib
import sys

__version__ = "0.1.0"

_MODULES = ("Elastic2D_CPML", "Elastic2D_sparse", "Elastic3D_SSG_CPML", "Elastic3D_SSG_SPML",
            "Elastic3D_SSG_Tensor")


def install_src_alias() -> None:
    """Register this package as ``src`` and its modules as ``src.<Module>``."""
    pkg = sys.modules[__name__]
    sys.modules.setdefault("src", pkg)
    for m in _MODULES:
        sys.modules.setdefault(f"src.{m}", importlib.import_module(f"{__name__}.{m}"))
