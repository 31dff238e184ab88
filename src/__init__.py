"""Drop-in ``src`` package: the reference drivers' imports
(``from src.Elastic3D_SSG_Tensor import ElasticWAVE``, e.g.
test3d_homegenous_tensor.py:4) resolve here, with the repository root on
``sys.path``, to the MI355X classes of ``seiswave_taichi_amd``.  The driver
scripts' own ``import taichi`` is theirs to satisfy (Taichi, or the host
stand-in that ``seiswave_taichi_amd.install_src_alias()`` registers)."""
