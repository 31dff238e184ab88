"""Host-side PML coefficient profiles (SetADEPML2D / SetADEPML3D).

Restates src/Elastic3D_SSG_CPML.py:324-522 (identical in
src/Elastic3D_SSG_Tensor.py:361-559, src/Elastic3D_SSG_SPML.py:408-606 and,
for two axes, src/Elastic2D_CPML.py:392-528):

* for each index i of an axis, at the integer coordinate ``x=(i-s)*d`` and the
  half coordinate ``x+d/2``, r is the normalised depth into the PML;
  ``d = d0*r**2`` with ``d0 = -3*vp_max*ln(R)/(2L)``, ``alpha = alpha_max*(1-r)``
  (clamped at 0), ``kappa = 1+(kappa_max-1)*r**2``; outside the PML
  d = alpha = 0 and kappa = 1;
* d, alpha and kappa are computed in f64 and stored as f32; ``b`` and ``a`` are
  then computed in f64 from the stored f32 values and stored as f32, ``a`` only
  where ``|d| > 1e-6``.

On top of the reference arrays this module derives what the GPU kernels need:

* ``in``/``in_h``: 1 where the reference kernel's f32 coordinate test
  (``x <= xmin_pml or x >= xmax_pml``, src/Elastic3D_SSG_CPML.py:171) holds
  for this axis;
* the slab ``[0, lo) U [hi, n)`` outside which the axis's CPML coefficients are
  the identity (a = 0, kappa = 1), so the reference's memory variables of this
  axis are exactly 0 there and need no storage.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

PML_KEYS = ("vp_max", "Rcoef", "theta", "alpha_max_pml", "kmax_pml")


@dataclass
class AxisPML:
    """The ten reference arrays of one axis plus the kernel-side extras."""
    d: np.ndarray
    d_h: np.ndarray
    alpha: np.ndarray
    alpha_h: np.ndarray
    k: np.ndarray
    k_h: np.ndarray
    b: np.ndarray
    b_h: np.ndarray
    a: np.ndarray
    a_h: np.ndarray
    min_pml: float
    max_pml: float
    inside: np.ndarray    # uint8, integer points
    inside_h: np.ndarray  # uint8, half points
    lo: int
    hi: int

    @property
    def width(self) -> int:
        return self.lo + (len(self.d) - self.hi)


def _profile(n, s, d, half, mn, mx, thick, d0, amax, kmax, use_lo, use_hi):
    """Per-index d, alpha, kappa as f32 (src/Elastic3D_SSG_CPML.py:366-408)."""
    x = np.arange(n, dtype=np.float64) - s
    x = x * d + half
    r = np.zeros(n)
    lo = (x <= mn) & bool(use_lo)
    hi = (x >= mx) & bool(use_hi) & ~lo
    r[lo] = (mn - x[lo]) / thick
    r[hi] = (x[hi] - mx) / thick
    sel = lo | hi
    dd = np.where(sel, d0 * r ** 2, 0.0)
    al = np.where(sel, amax * (1 - r), 0.0)
    kk = np.where(sel, 1 + (kmax - 1) * r ** 2, 1.0)
    D = dd.astype(np.float32)
    A = al.astype(np.float32)
    K = kk.astype(np.float32)
    A[A < 0] = 0
    return D, A, K


def _ab(D, A, K, dt, theta):
    """b and a from the stored f32 profiles (src/Elastic3D_SSG_CPML.py:500-522)."""
    D64, A64, K64 = (v.astype(np.float64) for v in (D, A, K))
    q = D64 / K64 + A64
    b = ((1 - (1 - theta) * dt * q) / (1 + theta * dt * q)).astype(np.float32)
    a = np.zeros_like(b)
    m = np.abs(D64) > 1e-6
    a[m] = ((-dt) * D64[m] / (K64[m] * K64[m]) / (1 + theta * dt * q[m])).astype(np.float32)
    return b, a


def kernel_coordinate(n: int, s: int, d: float, half: bool) -> np.ndarray:
    """The reference kernel's f32 coordinate ``(i-s)*d (+ d/2)`` for every i."""
    d32 = np.float32(d)
    x = (np.arange(n, dtype=np.int64) - s).astype(np.float32) * d32
    if half:
        x = x + np.float32(d32 / np.float32(2))
    return x


def axis_pml(n: int, s: int, d: float, dt: float, thick: float, vp_max: float, rcoef: float,
             alpha_max: float, kmax: float, theta: float, use_lo: bool, use_hi: bool) -> AxisPML:
    """One axis of SetADEPML3D plus the kernel flags and the memory slab."""
    if thick <= 0:
        raise ValueError(f"PML thickness must be positive, got {thick}")
    if rcoef <= 0:
        raise ValueError(f"Rcoef must be positive, got {rcoef}")
    xmax = d * (n - 1 - 2 * s)
    mn = 0 + thick
    mx = xmax - thick
    d0 = -3.0 * vp_max * math.log(rcoef) / (2.0 * thick)
    D, A, K = _profile(n, s, d, 0.0, mn, mx, thick, d0, alpha_max, kmax, use_lo, use_hi)
    Dh, Ah, Kh = _profile(n, s, d, d / 2, mn, mx, thick, d0, alpha_max, kmax, use_lo, use_hi)
    b, a = _ab(D, A, K, dt, theta)
    bh, ah = _ab(Dh, Ah, Kh, dt, theta)
    mn32, mx32 = np.float32(mn), np.float32(mx)
    xi = kernel_coordinate(n, s, d, False)
    xh = kernel_coordinate(n, s, d, True)
    inside = ((xi <= mn32) | (xi >= mx32)).astype(np.uint8)
    inside_h = ((xh <= mn32) | (xh >= mx32)).astype(np.uint8)
    lo, hi = memory_slab(a, k=K, a_h=ah, k_h=Kh, inside=inside, inside_h=inside_h)
    return AxisPML(D, Dh, A, Ah, K, Kh, b, bh, a, ah, mn, mx, inside, inside_h, lo, hi)


def memory_slab(a, k, a_h, k_h, inside, inside_h):
    """Slab bounds (lo, hi) of an axis for the CPML memory variables.

    An index needs memory storage when its integer or half coefficients are
    not the identity.  The reference only applies the CPML correction where the
    kernel's coordinate test holds; the slab form is exact as long as every
    non-identity coefficient sits at an index where that test holds, which is
    checked here (it fails only for a degenerate profile).
    """
    n = len(a)
    nontriv_i = (a != 0) | (k != 1)
    nontriv_h = (a_h != 0) | (k_h != 1)
    if np.any(nontriv_i & (inside == 0)) or np.any(nontriv_h & (inside_h == 0)):
        raise ValueError("PML profile has non-identity coefficients outside the reference's "
                         "PML coordinate test; the slab-confined CPML cannot reproduce it")
    nontriv = nontriv_i | nontriv_h
    mid = n // 2
    left = np.nonzero(nontriv[:mid])[0]
    right = np.nonzero(nontriv[mid:])[0]
    lo = int(left[-1]) + 1 if len(left) else 0
    hi = int(right[0]) + mid if len(right) else n
    if lo >= hi or np.any(nontriv[lo:hi]):
        lo, hi = n, n  # the whole axis is PML
    return lo, hi


def check_parameter(p: dict, keys) -> None:
    missing = [k for k in tuple(PML_KEYS) + tuple(keys) if k not in p]
    if missing:
        raise KeyError(f"PML parameter dict lacks {missing}")
