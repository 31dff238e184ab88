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
to these modules, ``src.BaseKernel`` and ``src.wigb`` to the host helpers of
those names (and ``import taichi`` to a host stand-in when Taichi is not
installed), so reference driver scripts run unchanged.
"""
import importlib
import sys

__version__ = "0.1.0"

_MODULES = ("Elastic2D_CPML", "Elastic2D_sparse", "Elastic3D_SSG_CPML", "Elastic3D_SSG_SPML",
            "Elastic3D_SSG_Tensor")
# host helpers the drivers import from src beside the solver classes
_HELPERS = ("BaseKernel", "wigb")


def install_src_alias(taichi_shim: bool = True) -> None:
    """Register this package as ``src`` and its modules as ``src.<Module>``;
    with ``taichi_shim``, also a host stand-in for ``taichi`` when Taichi is
    not installed (``taichi_shim.install``), so driver scripts that import it
    run unchanged."""
    if taichi_shim:
        from . import taichi_shim as _shim
        _shim.install()
    pkg = sys.modules[__name__]
    sys.modules.setdefault("src", pkg)
    for m in _MODULES + _HELPERS:
        sys.modules.setdefault(f"src.{m}", importlib.import_module(f"{__name__}.{m}"))
