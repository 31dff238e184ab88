"""CPU oracle driver (TEST INFRASTRUCTURE ONLY).

ctypes front end to ``liboracle_f32.so`` / ``liboracle_f64.so`` (built from
``seiswave_oracle.c`` by ``oracle/Makefile``), the plain-C restatement of the
SeisWave_Taichi reference kernels.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use this module, as the checker;
the product package never imports it.

Arrays use the reference layout: 3-D fields are ``(nx, ny, nz)`` C order (z
contiguous), 2-D fields ``(nx, nz)``.

Reference anchors (paths relative to the reference root):
  * moduli               src/Elastic3D_SSG_CPML.py:124-141 (NumPy, field dtype)
  * diff_coff            src/Elastic3D_SSG_CPML.py:311-322
  * SetADEPML3D / 2D     src/Elastic3D_SSG_CPML.py:324-522, src/Elastic2D_CPML.py:392-528
  * update_SSG           src/Elastic3D_SSG_CPML.py:144-309
  * update_SSG (Tensor)  src/Elastic3D_SSG_Tensor.py:166-345
  * update_SSG_SPML      src/Elastic3D_SSG_SPML.py:160-393
  * update_SSG_VS/SV/RSG src/Elastic2D_CPML.py:127-353
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")
MAX_ORDER = 8

_LIBS: dict = {}


def build() -> None:
    """Compile both oracle precisions with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(precision: str):
    if precision not in ("f32", "f64"):
        raise ValueError(f"oracle precision must be 'f32' or 'f64', got {precision!r}")
    if precision in _LIBS:
        return _LIBS[precision]
    path = os.path.join(_BUILD, f"liboracle_{precision}.so")
    if not os.path.exists(path):
        build()
    _LIBS[precision] = _bind(path, precision)
    return _LIBS[precision]


class library:
    """Context manager: oracle objects created inside use the oracle library at
    ``path`` (built from a variant of seiswave_oracle.c) in place of the
    committed build of that precision; used to check that the oracle's pins
    reject a deliberately mistaken restatement."""

    def __init__(self, path: str, precision: str = "f64"):
        self.path, self.precision = path, precision

    def __enter__(self):
        self._saved = _LIBS.get(self.precision)
        _lib(self.precision)  # the coefficient helpers keep using a real build
        _LIBS[self.precision] = _bind(self.path, self.precision)
        return self

    def __exit__(self, *exc):
        if self._saved is None:
            _LIBS.pop(self.precision, None)
        else:
            _LIBS[self.precision] = self._saved
        return False


def _bind(path: str, precision: str):
    lib = C.CDLL(path)
    real = C.c_float if precision == "f32" else C.c_double
    P = C.POINTER(real)

    class Axis(C.Structure):
        _fields_ = [(n, P) for n in ("d", "d_h", "alpha", "alpha_h", "k", "k_h", "b", "b_h", "a", "a_h")]

    class State3D(C.Structure):
        _fields_ = (
            [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("s", C.c_int32)]
            + [(n, real) for n in ("dx", "dy", "dz", "dt")]
            + [("c", real * MAX_ORDER)]
            + [(n, real) for n in ("xmin_pml", "xmax_pml", "ymin_pml", "ymax_pml", "zmin_pml", "zmax_pml")]
            + [("px", Axis), ("py", Axis), ("pz", Axis)]
            + [(n, P) for n in ("rho", "lam", "mu", "vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")]
            + [("mem", P * 18), ("split", P * 24)]
            + [("isx", C.c_int32), ("isy", C.c_int32), ("isz", C.c_int32), ("mt", real * 6)]
            + [("nrec", C.c_int32), ("rec", C.POINTER(C.c_int32)), ("data", P)]
        )

    class State2D(C.Structure):
        _fields_ = (
            [("nx", C.c_int32), ("nz", C.c_int32), ("s", C.c_int32)]
            + [(n, real) for n in ("dx", "dz", "dt")]
            + [("c", real * MAX_ORDER)]
            + [(n, real) for n in ("xmin_pml", "xmax_pml", "zmin_pml", "zmax_pml")]
            + [("px", Axis), ("pz", Axis)]
            + [(n, P) for n in ("rho", "lam", "mu", "vx", "vz", "sxx", "szz", "sxz")]
            + [("mem", P * 10)]
            + [(n, C.c_int32) for n in ("isx", "isz", "rsx", "rsz")]
            + [("data", P)]
        )

    lib.or_diff_coff.argtypes = [C.c_int32, C.POINTER(C.c_double)]
    lib.or_diff_coff.restype = C.c_int
    lib.or_ricker.argtypes = [C.c_int32, real, real, real]
    lib.or_ricker.restype = real
    lib.or_ricker2.argtypes = [C.c_int32, real, real, real, real]
    lib.or_ricker2.restype = real
    lib.or_pml_axis.argtypes = [C.c_int32, C.c_int32] + [C.c_double] * 8 + [
        C.c_int32, C.c_int32, C.POINTER(Axis), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    for name in ("or3d_cpml_step", "or3d_tensor_step", "or3d_spml_step"):
        getattr(lib, name).argtypes = [C.POINTER(State3D), C.c_int32, real]
    lib.or3d_cpml_velocity.argtypes = [C.POINTER(State3D), C.c_int32]
    for name in ("or3d_cpml_stress", "or3d_spml_velocity", "or3d_spml_stress"):
        getattr(lib, name).argtypes = [C.POINTER(State3D)]
    lib.or3d_aea_init.argtypes = [C.POINTER(State3D), P, P]
    for name in ("or2d_vs_step", "or2d_sv_step"):
        getattr(lib, name).argtypes = [C.POINTER(State2D), C.c_int32, real]
    lib.or2d_rsg_step.argtypes = [C.POINTER(State2D), real]
    lib.or2d_aea_init.argtypes = [C.POINTER(State2D)]
    lib.or2d_aea.argtypes = [C.POINTER(State2D)]
    lib.or_set_threads.argtypes = [C.c_int32]
    lib.or_set_threads.restype = C.c_int
    lib.Axis, lib.State3D, lib.State2D, lib.real, lib.P = Axis, State3D, State2D, real, P
    lib.np_dtype = np.float32 if precision == "f32" else np.float64
    return lib


def _ptr(lib, a: np.ndarray):
    assert a.flags.c_contiguous and a.dtype == lib.np_dtype
    return a.ctypes.data_as(lib.P)


def set_threads(n: int, precision: str = "f32") -> int:
    """OpenMP threads for the oracle loops; returns the count in effect."""
    return int(_lib(precision).or_set_threads(int(n)))


# ------------------------------------------------------------------ helpers
def diff_coff(order: int) -> np.ndarray:
    """f64 solution of the staggered Taylor system (src/Elastic3D_SSG_CPML.py:311-322)."""
    lib = _lib("f32")
    out = (C.c_double * order)()
    if lib.or_diff_coff(order, out) != 0:
        raise ValueError(f"order {order} outside 1..{MAX_ORDER}")
    return np.array(out[:], dtype=np.float64)


def ricker(nt: int, dt: float, freq: float, scale: float, precision: str = "f32") -> float:
    """src/BaseFun.py:4-11, evaluated in the oracle precision."""
    return float(_lib(precision).or_ricker(nt, dt, freq, scale))


def ricker2(nt: int, dt: float, t0: float, freq: float, scale: float, precision: str = "f32") -> float:
    """src/BaseFun.py:13-19, evaluated in the oracle precision."""
    return float(_lib(precision).or_ricker2(nt, dt, t0, freq, scale))


def pml_axis(n, s, d, dt, thick, vp_max, rcoef, alpha_max, kmax, theta, use_lo, use_hi,
             precision: str = "f32"):
    """One axis of SetADEPML3D (src/Elastic3D_SSG_CPML.py:364-522).

    Returns ``(arrays, min_pml, max_pml)`` with ``arrays`` a dict of the ten
    per-index arrays (d, d_h, alpha, alpha_h, k, k_h, b, b_h, a, a_h).
    """
    lib = _lib(precision)
    arrs = {name: np.zeros(n, dtype=lib.np_dtype) for name, _ in lib.Axis._fields_}
    ax = lib.Axis(**{name: _ptr(lib, a) for name, a in arrs.items()})
    mn, mx = C.c_double(), C.c_double()
    lib.or_pml_axis(n, s, d, dt, thick, vp_max, rcoef, alpha_max, kmax, theta,
                    int(bool(use_lo)), int(bool(use_hi)), C.byref(ax), C.byref(mn), C.byref(mx))
    return arrs, mn.value, mx.value


def moduli(vs: np.ndarray, vp: np.ndarray, rho: np.ndarray, dtype):
    """mu = rho*vs**2, lam = rho*(vp**2 - 2*vs**2) in the field dtype
    (src/Elastic3D_SSG_CPML.py:124-141)."""
    vs, vp, rho = (np.asarray(a, dtype=dtype) for a in (vs, vp, rho))
    mu = rho * vs ** 2
    lam = rho * (vp ** 2 - 2 * vs ** 2)
    return mu.astype(dtype), lam.astype(dtype)


# ---------------------------------------------------------------- 3-D oracle
class Oracle3D:
    """One reference ``ElasticWAVE`` 3-D instance on the CPU.

    ``kind`` selects the reference class: ``"cpml"`` (Elastic3D_SSG_CPML),
    ``"tensor"`` (Elastic3D_SSG_Tensor) or ``"spml"`` (Elastic3D_SSG_SPML).
    ``receivers`` is an ``(nrec, 3)`` integer array of (i, j, k) cells.
    """

    FIELDS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")

    def __init__(self, kind, vs, vp, rho, dx, dy, dz, dt, src, receivers, nt, accuracy,
                 freq=100.0, src_scale=1.0, mt=None, precision="f32"):
        if kind not in ("cpml", "tensor", "spml"):
            raise ValueError(kind)
        self.kind, self.precision = kind, precision
        self.lib = lib = _lib(precision)
        dtype = lib.np_dtype
        self.shape = tuple(int(v) for v in np.shape(vs))
        nx, ny, nz = self.shape
        self.rho = np.ascontiguousarray(rho, dtype=dtype).copy()
        self.mu, self.lam = moduli(vs, vp, rho, dtype)
        self.fields = {f: np.zeros(self.shape, dtype=dtype) for f in self.FIELDS}
        nmem = 18 if kind != "spml" else 0
        nsplit = 24 if kind == "spml" else 0
        self.mem = [np.zeros(self.shape, dtype=dtype) for _ in range(nmem)]
        self.split = [np.zeros(self.shape, dtype=dtype) for _ in range(nsplit)]
        rec = np.asarray(receivers, dtype=np.int32).reshape(-1, 3)
        self.rec = np.ascontiguousarray(rec)
        self.data = np.zeros((nt, len(rec)), dtype=dtype)
        self.nt, self.s = nt, accuracy
        self.dx, self.dy, self.dz, self.dt = dx, dy, dz, dt
        self.freq, self.src_scale = freq, src_scale
        self.src = tuple(int(v) for v in src)
        c64 = diff_coff(accuracy)
        # The reference stores c as f32 whatever the field type (Q8); the f64
        # oracle keeps the exact f64 coefficients.
        self.c = c64.astype(np.float32).astype(dtype) if precision == "f32" else c64
        self.mt = np.zeros(6, dtype=dtype) if mt is None else np.asarray(mt, dtype=dtype)
        self.pml = None
        st = lib.State3D()
        st.nx, st.ny, st.nz, st.s = nx, ny, nz, accuracy
        st.dx, st.dy, st.dz, st.dt = dx, dy, dz, dt
        for i, v in enumerate(self.c):
            st.c[i] = float(v)
        st.rho, st.lam, st.mu = (_ptr(lib, a) for a in (self.rho, self.lam, self.mu))
        for f in self.FIELDS:
            setattr(st, f, _ptr(lib, self.fields[f]))
        for i, a in enumerate(self.mem):
            st.mem[i] = _ptr(lib, a)
        for i, a in enumerate(self.split):
            st.split[i] = _ptr(lib, a)
        st.isx, st.isy, st.isz = self.src
        for i, v in enumerate(self.mt):
            st.mt[i] = float(v)
        st.nrec = len(rec)
        st.rec = self.rec.ctypes.data_as(C.POINTER(C.c_int32))
        st.data = _ptr(lib, self.data)
        self.st = st

    def set_pml(self, pml_surface, p: dict):
        """SetADEPML3D (src/Elastic3D_SSG_CPML.py:324-522)."""
        lib, st = self.lib, self.st
        n = self.shape
        self.pml = {}
        for ax, (nn, d, thick, lo, hi) in enumerate(zip(
                n, (self.dx, self.dy, self.dz),
                (p["pml_x_thick"], p["pml_y_thick"], p["pml_z_thick"]),
                pml_surface[0::2], pml_surface[1::2])):
            arrs, mn, mx = pml_axis(nn, self.s, d, self.dt, thick, p["vp_max"], p["Rcoef"],
                                    p["alpha_max_pml"], p["kmax_pml"], p["theta"], lo, hi,
                                    self.precision)
            name = "xyz"[ax]
            self.pml[name] = arrs
            setattr(st, f"p{name}", lib.Axis(**{k: _ptr(lib, a) for k, a in arrs.items()}))
            setattr(st, f"{name}min_pml", mn)
            setattr(st, f"{name}max_pml", mx)

    def aea_init(self):
        """Elastic3D_SSG_Tensor.AEA_init (src/Elastic3D_SSG_Tensor.py:151-157)."""
        self.lib.or3d_aea_init(C.byref(self.st), _ptr(self.lib, self.rho), _ptr(self.lib, self.lam))

    def aea(self):
        """Elastic3D_SSG_Tensor.AEA (src/Elastic3D_SSG_Tensor.py:159-163):
        szz[i, j, accuracy] = 0 for every i, j."""
        self.fields["szz"][:, :, self.s] = 0

    def wavelet(self, it: int) -> float:
        """The reference's per-step source amplitude for this class."""
        if self.kind == "cpml":    # src/Elastic3D_SSG_CPML.py:245
            return ricker(it, self.dt, self.freq, 1000.0, self.precision)
        if self.kind == "spml":    # src/Elastic3D_SSG_SPML.py:294
            return ricker(it, self.dt, self.freq, self.src_scale, self.precision)
        return ricker2(it, self.dt, 0.03, self.freq, self.src_scale, self.precision)  # Tensor :186

    def step(self, it: int, amp: float | None = None):
        if self.pml is None:
            raise RuntimeError("SetADEPML3D must run before the first step")
        if amp is None:
            amp = self.wavelet(it)
        fn = {"cpml": self.lib.or3d_cpml_step, "tensor": self.lib.or3d_tensor_step,
              "spml": self.lib.or3d_spml_step}[self.kind]
        fn(C.byref(self.st), it, amp)

    def run(self, nsteps: int, it0: int = 0, amps=None):
        for n in range(nsteps):
            self.step(it0 + n, None if amps is None else float(amps[it0 + n]))

    def step_moment(self, it: int, amp: float):
        """One step of this class with the moment-tensor source of
        src/Elastic3D_SSG_Tensor.py:186-193 in place of the class's own
        source: s_ab[is] += amp*(-M_ab) before the velocity loops, then the
        class's velocity and stress loops and the vx record (the order
        sw_ssg3d_run uses for a stress source in every class; the reference
        has a moment source only in the Tensor class)."""
        if self.pml is None:
            raise RuntimeError("SetADEPML3D must run before the first step")
        if self.kind == "tensor":
            return self.step(it, amp)
        dt = self.lib.np_dtype
        a = dt(amp)
        for name, m in zip(("sxx", "syy", "szz", "sxy", "sxz", "syz"), self.mt):
            self.fields[name][self.src] += a * (-m)
        self.velocity()
        self.stress()
        for r, (i, j, k) in enumerate(self.rec):
            self.data[it, r] = self.fields["vx"][i, j, k]

    def velocity(self):
        """The velocity loops alone (no source, no record)."""
        if self.kind == "spml":
            self.lib.or3d_spml_velocity(C.byref(self.st))
        else:
            self.lib.or3d_cpml_velocity(C.byref(self.st), int(self.kind == "tensor"))

    def stress(self):
        """The stress loops alone."""
        if self.kind == "spml":
            self.lib.or3d_spml_stress(C.byref(self.st))
        else:
            self.lib.or3d_cpml_stress(C.byref(self.st))


# ---------------------------------------------------------------- 2-D oracle
class Oracle2D:
    """One reference ``Elastic2D_CPML.ElasticWAVE`` instance on the CPU."""

    FIELDS = ("vx", "vz", "sxx", "szz", "sxz")

    def __init__(self, vs, vp, rho, dx, dz, dt, isx, isz, rsx, rsz, nt, accuracy=3, freq=100.0,
                 precision="f32"):
        self.precision = precision
        self.lib = lib = _lib(precision)
        dtype = lib.np_dtype
        self.shape = tuple(int(v) for v in np.shape(vs))
        nx, nz = self.shape
        self.rho = np.ascontiguousarray(rho, dtype=dtype).copy()
        self.mu, self.lam = moduli(vs, vp, rho, dtype)
        self.fields = {f: np.zeros(self.shape, dtype=dtype) for f in self.FIELDS}
        self.mem = [np.zeros(self.shape, dtype=dtype) for _ in range(10)]
        self.data = np.zeros(nt, dtype=dtype)
        self.nt, self.s, self.dx, self.dz, self.dt, self.freq = nt, accuracy, dx, dz, dt, freq
        c64 = diff_coff(accuracy)
        self.c = c64.astype(np.float32).astype(dtype) if precision == "f32" else c64
        self.pml = None
        st = lib.State2D()
        st.nx, st.nz, st.s = nx, nz, accuracy
        st.dx, st.dz, st.dt = dx, dz, dt
        for i, v in enumerate(self.c):
            st.c[i] = float(v)
        st.rho, st.lam, st.mu = (_ptr(lib, a) for a in (self.rho, self.lam, self.mu))
        for f in self.FIELDS:
            setattr(st, f, _ptr(lib, self.fields[f]))
        for i, a in enumerate(self.mem):
            st.mem[i] = _ptr(lib, a)
        st.isx, st.isz, st.rsx, st.rsz = int(isx), int(isz), int(rsx), int(rsz)
        st.data = _ptr(lib, self.data)
        self.st = st

    def set_pml(self, pml_surface, p: dict):
        """SetADEPML2D (src/Elastic2D_CPML.py:392-528)."""
        lib, st = self.lib, self.st
        self.pml = {}
        for name, nn, d, thick, lo, hi in (
                ("x", self.shape[0], self.dx, p["pml_x_thick"], pml_surface[0], pml_surface[1]),
                ("z", self.shape[1], self.dz, p["pml_z_thick"], pml_surface[2], pml_surface[3])):
            arrs, mn, mx = pml_axis(nn, self.s, d, self.dt, thick, p["vp_max"], p["Rcoef"],
                                    p["alpha_max_pml"], p["kmax_pml"], p["theta"], lo, hi,
                                    self.precision)
            self.pml[name] = arrs
            setattr(st, f"p{name}", lib.Axis(**{k: _ptr(lib, a) for k, a in arrs.items()}))
            setattr(st, f"{name}min_pml", mn)
            setattr(st, f"{name}max_pml", mx)

    def vs_step(self, it, amp=None):
        """update_SSG_VS; amplitude Ricker2(t0=1.5/f0, scale 1) (src/Elastic2D_CPML.py:142,170)."""
        self._need_pml()
        if amp is None:
            amp = ricker2(it, self.dt, 1.5 / self.freq, self.freq, 1.0, self.precision)
        self.lib.or2d_vs_step(C.byref(self.st), it, amp)

    def sv_step(self, it, amp=None):
        """update_SSG_SV; amplitude Ricker(scale 1) (src/Elastic2D_CPML.py:245)."""
        self._need_pml()
        if amp is None:
            amp = ricker(it, self.dt, self.freq, 1.0, self.precision)
        self.lib.or2d_sv_step(C.byref(self.st), it, amp)

    def rsg_step(self, source):
        self._need_pml()
        self.lib.or2d_rsg_step(C.byref(self.st), source)

    def aea_init(self):
        self.lib.or2d_aea_init(C.byref(self.st))

    def aea(self):
        self.lib.or2d_aea(C.byref(self.st))

    def _need_pml(self):
        if self.pml is None:
            raise RuntimeError("SetADEPML2D must run before the first step")
