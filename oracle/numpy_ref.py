"""Independent vectorised NumPy restatement of the reference step (TEST
INFRASTRUCTURE ONLY).

Written separately from ``seiswave_oracle.c`` (array slicing instead of
per-cell loops, masks instead of branches) so that the two can cross-check
each other on tiny grids in f64.  Reference layout (nx, ny, nz).  Anchors:
src/Elastic3D_SSG_CPML.py:144-309, src/Elastic3D_SSG_Tensor.py:166-345,
src/Elastic3D_SSG_SPML.py:160-393, src/Elastic2D_CPML.py:127-353.
"""
from __future__ import annotations

import numpy as np


def _v(f, rng, off):
    return f[tuple(slice(r[0] + o, r[1] + o) for r, o in zip(rng, off))]


def _unit(nd, axis, v):
    o = [0] * nd
    o[axis] = v
    return o


def dfm(f, axis, rng, c):
    nd = f.ndim
    acc = 0.0
    for i, ci in enumerate(c):
        acc = acc + ci * (_v(f, rng, _unit(nd, axis, i + 1)) - _v(f, rng, _unit(nd, axis, -i)))
    return acc


def dbm(f, axis, rng, c):
    nd = f.ndim
    acc = 0.0
    for i, ci in enumerate(c):
        acc = acc + ci * (_v(f, rng, _unit(nd, axis, i)) - _v(f, rng, _unit(nd, axis, -i - 1)))
    return acc


def _coord(rng_ax, s, d, half):
    i = np.arange(rng_ax[0], rng_ax[1], dtype=np.float64)
    return (i - s) * d + (d / 2 if half else 0.0)


def _mask(rng, s, ds, halves, bounds):
    """In-PML mask: any axis coordinate <= min_pml or >= max_pml."""
    nd = len(rng)
    m = None
    for ax in range(nd):
        x = _coord(rng[ax], s, ds[ax], halves[ax])
        mn, mx = bounds[ax]
        mm = (x <= mn) | (x >= mx)
        shape = [1] * nd
        shape[ax] = len(x)
        mm = mm.reshape(shape)
        m = mm if m is None else (m | mm)
    return np.broadcast_to(m, tuple(r[1] - r[0] for r in rng))


def _coef(arr, ax, rng, nd):
    a = arr[rng[ax][0]:rng[ax][1]]
    shape = [1] * nd
    shape[ax] = len(a)
    return a.reshape(shape)


class NumpyRef3D:
    """fp64 restatement of the 3-D classes ("cpml", "tensor", "spml")."""

    def __init__(self, kind, rho, mu, lam, dx, dy, dz, dt, c, pml, bounds, src, mt=None, rec=()):
        self.kind = kind
        self.rho, self.mu, self.lam = (np.asarray(a, np.float64) for a in (rho, mu, lam))
        self.shape = self.rho.shape
        self.d = (dx, dy, dz)
        self.dt = dt
        self.c = np.asarray(c, np.float64)
        self.s = len(c)
        self.pml = pml          # per axis dict of arrays (a, b, k, a_h, ..., d, d_h)
        self.bounds = bounds    # per axis (min_pml, max_pml)
        self.src = src
        self.mt = np.zeros(6) if mt is None else np.asarray(mt, np.float64)
        self.rec = list(rec)
        z = lambda: np.zeros(self.shape)  # noqa: E731
        self.f = {n: z() for n in ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")}
        self.m = {}
        self.q = {}
        self.data = []

    # CPML correction of one derivative on a range
    def _cpml(self, D, key, ax, half, rng, mask):
        nd = 3
        P = self.pml[ax]
        sfx = "_h" if half else ""
        a, b, k = (_coef(P[n + sfx], ax, rng, nd) for n in ("a", "b", "k"))
        m = self.m.setdefault(key, np.zeros(self.shape))
        mv = _v(m, rng, (0, 0, 0))
        new = b * mv + a * D
        mv[mask] = np.broadcast_to(new, mask.shape)[mask]
        out = D / k + mv
        return np.where(mask, out, D)

    def _rng(self, lo, hi):
        n = self.shape
        return tuple((lo[a], n[a] + hi[a]) for a in range(3))

    def velocity(self):
        s, c, f, (dx, dy, dz), dt = self.s, self.c, self.f, self.d, self.dt
        rho = self.rho
        # vx
        r = self._rng((s, s + 1, s + 1), (-s - 1, -s, -s))
        mk = _mask(r, s, self.d, (True, False, False), self.bounds)
        A = dfm(f["sxx"], 0, r, c) / dx
        B = dbm(f["sxy"], 1, r, c) / dy
        Cc = dbm(f["sxz"], 2, r, c) / dz
        if self.kind == "spml":
            self._split_v("vx", r, mk, (A, B, Cc), (True, False, False), 1.0 / _v(rho, r, (0, 0, 0)),
                          order=(0, 1, 2), dt_first=False)
        else:
            A = self._cpml(A, "sxx_dx", 0, True, r, mk)
            B = self._cpml(B, "sxy_dy", 1, False, r, mk)
            Cc = self._cpml(Cc, "sxz_dz", 2, False, r, mk)
            _v(f["vx"], r, (0, 0, 0))[...] += (A + Cc + B) * dt / _v(rho, r, (0, 0, 0))
        # vy
        r = self._rng((s + 1, s, s + 1), (-s, -s - 1, -s))
        mk = _mask(r, s, self.d, (False, True, False), self.bounds)
        rh = 0.5 * (_v(rho, r, (0, 0, 0)) + _v(rho, r, (0, 1, 0)))
        A = dbm(f["sxy"], 0, r, c) / dx
        B = dfm(f["syy"], 1, r, c) / dy
        Cc = dbm(f["syz"], 2, r, c) / dz
        if self.kind == "spml":
            self._split_v("vy", r, mk, (A, B, Cc), (False, True, False), 1.0 / rh, order=(0, 1, 2), dt_first=True)
        else:
            A = self._cpml(A, "sxy_dx", 0, False, r, mk)
            B = self._cpml(B, "syy_dy", 1, True, r, mk)
            Cc = self._cpml(Cc, "syz_dz", 2, False, r, mk)
            _v(f["vy"], r, (0, 0, 0))[...] += (A + B + Cc) * dt / rh
        # vz
        r = self._rng((s + 1, s + 1, s), (-s, -s, -s - 1))
        mk = _mask(r, s, self.d, (False, False, True), self.bounds)
        rh = 0.25 * (_v(rho, r, (0, 0, 0)) + _v(rho, r, (0, 1, 0)) + _v(rho, r, (1, 1, 0)) + _v(rho, r, (1, 0, 0)))
        A = dbm(f["sxz"], 0, r, c) / dx
        B = dbm(f["syz"], 1, r, c) / dy
        Cc = dfm(f["szz"], 2, r, c) / dz
        if self.kind == "spml":
            self._split_v("vz", r, mk, (A, B, Cc), (False, False, True), 1.0 / rh, order=(0, 1, 2), dt_first=True)
        else:
            A = self._cpml(A, "sxz_dx", 0, False, r, mk)
            B = self._cpml(B, "syz_dy", 1, False, r, mk)
            Cc = self._cpml(Cc, "szz_dz", 2, True, r, mk)
            tot = (A + B + Cc) if self.kind == "tensor" else (Cc + A + B)
            _v(f["vz"], r, (0, 0, 0))[...] += tot * dt / rh

    def _split_v(self, name, r, mk, D, halves, rho_over, order, dt_first):
        dt = self.dt
        tot = None
        for ax in order:
            P = self.pml[ax]
            dd = _coef(P["d_h" if halves[ax] else "d"], ax, r, 3)
            key = f"{name}_{'xyz'[ax]}"
            q = self.q.setdefault(key, np.zeros(self.shape))
            qv = _v(q, r, (0, 0, 0))
            incr = (dt * D[ax] * rho_over) if dt_first else (D[ax] * dt * rho_over)
            new = ((1 - 0.5 * dt * dd) * qv + incr) / (1 + 0.5 * dt * dd)
            qv[mk] = np.broadcast_to(new, mk.shape)[mk]
            tot = qv.copy() if tot is None else tot + qv
        fv = _v(self.f[name], r, (0, 0, 0))
        plain = fv + (D[0] + D[1] + D[2]) * dt * rho_over
        fv[...] = np.where(mk, tot, plain)

    def _split_s(self, name, r, mk, terms, dirs, halves):
        """terms: per split direction the increment (already multiplied by dt)."""
        dt = self.dt
        tot = None
        for t, ax in zip(terms, dirs):
            P = self.pml[ax]
            dd = _coef(P["d_h" if halves[ax] else "d"], ax, r, 3)
            key = f"{name}_{'xyz'[ax]}"
            q = self.q.setdefault(key, np.zeros(self.shape))
            qv = _v(q, r, (0, 0, 0))
            new = ((1 - 0.5 * dt * dd) * qv + t) / (1 + 0.5 * dt * dd)
            qv[mk] = np.broadcast_to(new, mk.shape)[mk]
            tot = qv.copy() if tot is None else tot + qv
        return tot

    def stress(self):
        s, c, f, (dx, dy, dz), dt = self.s, self.c, self.f, self.d, self.dt
        mu, lam = self.mu, self.lam
        # normal
        r = self._rng((s + 1, s + 1, s + 1), (-s, -s, -s))
        mk = _mask(r, s, self.d, (False, False, False), self.bounds)
        A = dbm(f["vx"], 0, r, c) / dx
        B = dbm(f["vy"], 1, r, c) / dy
        Cc = dbm(f["vz"], 2, r, c) / dz
        if self.kind == "spml":
            la, m_ = _v(lam, r, (0, 0, 0)), _v(mu, r, (0, 0, 0))
            l2 = la + 2 * m_
            plain = {
                "sxx": _v(f["sxx"], r, (0, 0, 0)) + (l2 * A + la * (Cc + B)) * dt,
                "syy": _v(f["syy"], r, (0, 0, 0)) + (l2 * B + la * (A + Cc)) * dt,
                "szz": _v(f["szz"], r, (0, 0, 0)) + (l2 * Cc + la * (A + B)) * dt,
            }
            ins = {
                "sxx": self._split_s("sxx", r, mk, (dt * l2 * A, dt * la * B, dt * la * Cc), (0, 1, 2), (0, 0, 0)),
                "syy": self._split_s("syy", r, mk, (dt * la * A, dt * l2 * B, dt * la * Cc), (0, 1, 2), (0, 0, 0)),
                "szz": self._split_s("szz", r, mk, (dt * la * A, dt * la * B, dt * l2 * Cc), (0, 1, 2), (0, 0, 0)),
            }
            for n in ("sxx", "syy", "szz"):
                _v(f[n], r, (0, 0, 0))[...] = np.where(mk, ins[n], plain[n])
        else:
            lh = 0.5 * (_v(lam, r, (0, 0, 0)) + _v(lam, r, (1, 0, 0)))
            mh = 0.5 * (_v(mu, r, (0, 0, 0)) + _v(mu, r, (1, 0, 0)))
            lpm = lh + 2 * mh
            A = self._cpml(A, "dvx_dx", 0, False, r, mk)
            B = self._cpml(B, "dvy_dy", 1, False, r, mk)
            Cc = self._cpml(Cc, "dvz_dz", 2, False, r, mk)
            _v(f["sxx"], r, (0, 0, 0))[...] += (lpm * A + lh * (Cc + B)) * dt
            _v(f["syy"], r, (0, 0, 0))[...] += (lpm * B + lh * (A + Cc)) * dt
            _v(f["szz"], r, (0, 0, 0))[...] += (lpm * Cc + lh * (A + B)) * dt
        # sxy
        r = self._rng((s + 1, s + 1, s), (-s, -s, -s))
        mk = _mask(r, s, self.d, (True, True, False), self.bounds)
        ma = 0.25 * (_v(mu, r, (0, 1, 0)) + _v(mu, r, (0, 0, 0)) + _v(mu, r, (0, 0, 0)) + _v(mu, r, (1, 0, 0)))
        A = dfm(f["vy"], 0, r, c) / dx
        B = dfm(f["vx"], 1, r, c) / dy
        self._shear("sxy", r, mk, ma, A, B, (0, 1), (True, True), ("dvy_dx", "dvx_dy"))
        # sxz
        r = self._rng((s, s, s), (-s - 1, -s, -s - 1))
        mk = _mask(r, s, self.d, (True, False, True), self.bounds)
        ma = 0.25 * (_v(mu, r, (0, 0, 0)) + _v(mu, r, (1, 0, 0)) + _v(mu, r, (0, 0, 0)) + _v(mu, r, (0, 1, 0)))
        A = dfm(f["vz"], 0, r, c) / dx
        B = dfm(f["vx"], 2, r, c) / dz
        self._shear("sxz", r, mk, ma, A, B, (0, 2), (True, True), ("dvz_dx", "dvx_dz"))
        # syz
        r = self._rng((s, s, s), (-s, -s - 1, -s - 1))
        mk = _mask(r, s, self.d, (False, True, True), self.bounds)
        ma = 0.25 * (_v(mu, r, (0, 1, 0)) + _v(mu, r, (0, 0, 0)) + _v(mu, r, (0, 0, 1)) + _v(mu, r, (0, 0, 0)))
        A = dfm(f["vz"], 1, r, c) / dy
        B = dfm(f["vy"], 2, r, c) / dz
        self._shear("syz", r, mk, ma, A, B, (1, 2), (True, True), ("dvz_dy", "dvy_dz"))

    def _shear(self, name, r, mk, ma, A, B, axes, halves, keys):
        dt = self.dt
        fv = _v(self.f[name], r, (0, 0, 0))
        if self.kind == "spml":
            hv = [False, False, False]
            for ax, h in zip(axes, halves):
                hv[ax] = h
            ins = self._split_s(name, r, mk, (dt * ma * A, dt * ma * B), axes, hv)
            plain = fv + ma * (B + A) * dt
            fv[...] = np.where(mk, ins, plain)
        else:
            A = self._cpml(A, keys[0], axes[0], halves[0], r, mk)
            B = self._cpml(B, keys[1], axes[1], halves[1], r, mk)
            fv[...] += ma * (B + A) * dt

    def step(self, amp):
        f, dt = self.f, self.dt
        if self.kind == "tensor":
            i, j, k = self.src
            for n, m in zip(("sxx", "syy", "szz", "sxy", "sxz", "syz"), self.mt):
                f[n][i, j, k] += amp * (-m)
        self.velocity()
        if self.kind != "tensor":
            i, j, k = self.src
            f["vz"][i, j, k] += dt * amp / self.rho[i, j, k]
        self.stress()
        self.data.append([f["vx"][r] for r in self.rec])


class NumpyRef2D:
    """fp64 restatement of the 2-D class (schemes "VS", "SV", "RSG")."""

    def __init__(self, scheme, rho, mu, lam, dx, dz, dt, c, pml, bounds, src, rec):
        self.scheme = scheme
        self.rho, self.mu, self.lam = (np.asarray(a, np.float64) for a in (rho, mu, lam))
        self.shape = self.rho.shape
        self.d = (dx, dz)
        self.dt = dt
        self.c = np.asarray(c, np.float64)
        self.s = len(c)
        self.pml, self.bounds, self.src, self.rec = pml, bounds, src, rec
        self.f = {n: np.zeros(self.shape) for n in ("vx", "vz", "sxx", "szz", "sxz")}
        self.m = {}
        self.data = []

    def _rng(self, lo, hi):
        return tuple((lo[a], self.shape[a] + hi[a]) for a in range(2))

    def _cpml(self, D, key, ax, half, rng, mask):
        P = self.pml[ax]
        sfx = "_h" if half else ""
        a, b, k = (_coef(P[n + sfx], ax, rng, 2) for n in ("a", "b", "k"))
        m = self.m.setdefault(key, np.zeros(self.shape))
        mv = _v(m, rng, (0, 0))
        new = b * mv + a * D
        mv[mask] = np.broadcast_to(new, mask.shape)[mask]
        return np.where(mask, D / k + mv, D)

    def _rot(self, f, r, fwd, kind):
        """Rotated differences (src/Differential2D.py:32-56)."""
        acc = 0.0
        for i, ci in enumerate(self.c):
            if kind == "x":
                a, b = ((i + 1, i + 1), (-i, -i)) if fwd else ((i, i), (-i - 1, -i - 1))
            else:
                a, b = ((-i, i + 1), (i + 1, -i)) if fwd else ((-i - 1, i), (i, -i - 1))
            acc = acc + ci * (_v(f, r, a) - _v(f, r, b))
        return acc

    def step(self, amp):
        s, c, f, (dx, dz), dt = self.s, self.c, self.f, self.d, self.dt
        rho, mu, lam = self.rho, self.mu, self.lam
        sc = self.scheme
        if sc == "VS":
            r = self._rng((s + 1, s + 1), (-s, -s))
            mk = _mask(r, s, self.d, (False, False), self.bounds)
            A = self._cpml(dbm(f["sxx"], 0, r, c) / dx, "sxx_dx", 0, False, r, mk)
            B = self._cpml(dbm(f["sxz"], 1, r, c) / dz, "sxz_dz", 1, False, r, mk)
            _v(f["vx"], r, (0, 0))[...] += (A + B) * dt / _v(rho, r, (0, 0))
            r = self._rng((s, s), (-s - 1, -s - 1))
            mk = _mask(r, s, self.d, (True, True), self.bounds)
            rh = 0.25 * (_v(rho, r, (0, 0)) + _v(rho, r, (0, 1)) + _v(rho, r, (1, 1)) + _v(rho, r, (1, 0)))
            A = self._cpml(dfm(f["sxz"], 0, r, c) / dx, "sxz_dx", 0, True, r, mk)
            B = self._cpml(dfm(f["szz"], 1, r, c) / dz, "szz_dz", 1, True, r, mk)
            _v(f["vz"], r, (0, 0))[...] += (B + A) * dt / rh
            i, j = self.src
            f["vz"][i, j] += dt * amp / rho[i, j]
            r = self._rng((s, s + 1), (-s - 1, -s))
            mk = _mask(r, s, self.d, (True, False), self.bounds)
            lh = 0.5 * (_v(lam, r, (0, 0)) + _v(lam, r, (1, 0)))
            mh = 0.5 * (_v(mu, r, (0, 0)) + _v(mu, r, (1, 0)))
            lpm = lh + 2 * mh
            A = self._cpml(dfm(f["vx"], 0, r, c) / dx, "dvx_dx", 0, True, r, mk)
            B = self._cpml(dbm(f["vz"], 1, r, c) / dz, "dvz_dz", 1, False, r, mk)
            _v(f["sxx"], r, (0, 0))[...] += (lpm * A + lh * B) * dt
            _v(f["szz"], r, (0, 0))[...] += (lpm * B + lh * A) * dt
            r = self._rng((s + 1, s), (-s, -s - 1))
            mk = _mask(r, s, self.d, (False, True), self.bounds)
            mz = 0.5 * (_v(mu, r, (0, 1)) + _v(mu, r, (0, 0)))
            A = self._cpml(dbm(f["vz"], 0, r, c) / dx, "dvz_dx", 0, False, r, mk)
            B = self._cpml(dfm(f["vx"], 1, r, c) / dz, "dvx_dz", 1, True, r, mk)
            _v(f["sxz"], r, (0, 0))[...] += mz * (B + A) * dt
            self.data.append(f["vz"][self.rec])
        elif sc == "SV":
            r = self._rng((s, s + 1), (-s - 1, -s))
            mk = _mask(r, s, self.d, (True, False), self.bounds)
            rh = (_v(rho, r, (0, 0)) + _v(rho, r, (1, 0))) / 2
            A = self._cpml(dfm(f["sxx"], 0, r, c) / dx, "sxx_dx", 0, True, r, mk)
            B = self._cpml(dbm(f["sxz"], 1, r, c) / dz, "sxz_dz", 1, False, r, mk)
            _v(f["vx"], r, (0, 0))[...] += (A + B) * dt / rh
            r = self._rng((s + 1, s), (-s, -s - 1))
            mk = _mask(r, s, self.d, (False, True), self.bounds)
            rh = 0.5 * (_v(rho, r, (0, 0)) + _v(rho, r, (0, 1)))
            A = self._cpml(dbm(f["sxz"], 0, r, c) / dx, "sxz_dx", 0, False, r, mk)
            B = self._cpml(dfm(f["szz"], 1, r, c) / dz, "szz_dz", 1, True, r, mk)
            _v(f["vz"], r, (0, 0))[...] += (B + A) * dt / rh
            i, j = self.src
            f["vx"][i, j] += dt * amp / rho[i, j]
            r = self._rng((s + 1, s + 1), (-s, -s))
            mk = _mask(r, s, self.d, (False, False), self.bounds)
            la = _v(lam, r, (0, 0))
            lpm = la + 2 * _v(mu, r, (0, 0))
            A = self._cpml(dbm(f["vx"], 0, r, c) / dx, "dvx_dx", 0, False, r, mk)
            B = self._cpml(dbm(f["vz"], 1, r, c) / dz, "dvz_dz", 1, False, r, mk)
            _v(f["sxx"], r, (0, 0))[...] += (lpm * A + la * B) * dt
            _v(f["szz"], r, (0, 0))[...] += (lpm * B + la * A) * dt
            r = self._rng((s + 1, s + 1), (-s, -s))
            mk = _mask(r, s, self.d, (True, True), self.bounds)
            ma = 0.25 * (_v(mu, r, (0, 1)) + _v(mu, r, (0, 0)) + _v(mu, r, (1, 1)) + _v(mu, r, (1, 0)))
            A = self._cpml(dfm(f["vz"], 0, r, c) / dx, "dvz_dx", 0, True, r, mk)
            B = self._cpml(dfm(f["vx"], 1, r, c) / dz, "dvx_dz", 1, True, r, mk)
            _v(f["sxz"], r, (0, 0))[...] += ma * (B + A) * dt
            self.data.append(f["vx"][self.rec])
        else:  # RSG
            r = self._rng((s + 1, s + 1), (-s, -s))
            mk = _mask(r, s, self.d, (False, False), self.bounds)
            xrx, xrz = self._rot(f["sxx"], r, False, "x"), self._rot(f["sxx"], r, False, "z")
            zrx, zrz = self._rot(f["szz"], r, False, "x"), self._rot(f["szz"], r, False, "z")
            crx, crz = self._rot(f["sxz"], r, False, "x"), self._rot(f["sxz"], r, False, "z")
            A = self._cpml((xrz - xrx) / (2 * dx), "sxx_dx", 0, False, r, mk)
            B = self._cpml((zrz + zrx) / (2 * dz), "szz_dz", 1, False, r, mk)
            Cx = self._cpml((crz - crx) / (2 * dx), "sxz_dx", 0, False, r, mk)
            Cz = self._cpml((crz + crx) / (2 * dz), "sxz_dz", 1, False, r, mk)
            rr = _v(rho, r, (0, 0))
            _v(f["vx"], r, (0, 0))[...] += (A + Cz) * dt / rr
            _v(f["vz"], r, (0, 0))[...] += (B + Cx) * dt / rr
            i, j = self.src
            for a in (-1, 0):
                for b in (-1, 0):
                    f["vz"][i + a, j + b] += dt * amp / (8 * rho[i + a, j + b])
            r = self._rng((s, s), (-s - 1, -s - 1))
            mk = _mask(r, s, self.d, (True, True), self.bounds)
            la, m_ = _v(lam, r, (0, 0)), _v(mu, r, (0, 0))
            lpm = la + 2 * m_
            vxrx, vxrz = self._rot(f["vx"], r, True, "x"), self._rot(f["vx"], r, True, "z")
            vzrx, vzrz = self._rot(f["vz"], r, True, "x"), self._rot(f["vz"], r, True, "z")
            dvxdx = self._cpml((vxrz - vxrx) / (2 * dx), "dvx_dx", 0, True, r, mk)
            dvxdz = self._cpml((vxrz + vxrx) / (2 * dz), "dvx_dz", 1, True, r, mk)
            dvzdx = self._cpml((vzrz - vzrx) / (2 * dx), "dvz_dx", 0, True, r, mk)
            dvzdz = self._cpml((vzrz + vzrx) / (2 * dz), "dvz_dz", 1, True, r, mk)
            _v(f["sxx"], r, (0, 0))[...] += (lpm * dvxdx + la * dvzdz) * dt
            _v(f["szz"], r, (0, 0))[...] += (lpm * dvzdz + la * dvxdx) * dt
            _v(f["sxz"], r, (0, 0))[...] += (m_ * (dvxdz + dvzdx)) * dt
