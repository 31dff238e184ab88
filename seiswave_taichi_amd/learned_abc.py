"""Learned thin absorbing boundary (no counterpart in the reference).

A thin CPML (4 cells) plus one small MLP per face that corrects the fields of
the 4 outermost interior cells after every step, replacing most of a wide
(20-cell) CPML.  The networks run in the fused HIP kernel sw_band_correct;
training uses the HIP dense forward/backward/Adam kernels (sw_mlp_*), on
samples generated by the plain solver itself.

Training targets (teacher forcing).  A wide-CPML run of a random model is the
reference.  A thin-CPML run of the same interior is reset to the reference's
interior state before every step; after the step, the features are the thin
run's band (9 fields x 4 cells, "cur") and its change over the step (cur minus
the band before the step), and the target is the reference band minus cur:
the one-step error the thin boundary makes next to itself.  At inference the
corrected band of the previous step takes the place of "the band before the
step", as in training (with a perfect correction the two coincide).

Scale.  The wave equation is linear, so the correction is made positively
homogeneous: each column's features are divided by their RMS s (after a fixed
per-channel scaling) before the network and its output is multiplied by s.
Training samples are drawn with probability proportional to s^2, so the loss
is the raw squared one-step error.

The band carries all 9 field components of the 3-D solver, so the training
runs are small 3-D boxes (layered, perturbed models, random moment-tensor
sources) and the evaluation is a 200^3 homogeneous case against a 20-cell
CPML.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import torch

from . import _lib as L
from .coeffs import ricker2

DEPTH, NFEAT, NOUT, NH = 4, 72, 36, 64
FACES = ("x-lo", "x-hi", "y-lo", "y-hi", "z-lo", "z-hi")


SwBand = L.SwBand


def _sig():
    return L.load()


def param_count() -> int:
    return int(_sig().sw_mlp_param_count())


def layout():
    """Offsets of (W1, b1, W2, b2, W3, b3, in_scale, out_scale) in one face's parameters."""
    sizes = [NFEAT * NH, NH, NH * NH, NH, NH * NOUT, NOUT, NFEAT, NOUT]
    off = np.concatenate([[0], np.cumsum(sizes)])
    names = ("W1", "b1", "W2", "b2", "W3", "b3", "in_scale", "out_scale")
    return {n: (int(off[i]), int(off[i + 1])) for i, n in enumerate(names)}


def band_of(engine) -> SwBand:
    """Interior box (outside every CPML slab) of a 3-D engine."""
    b = SwBand()
    (ax, ay, az) = engine.pml
    b.ox, b.oy, b.oz = ax.lo, ay.lo, az.lo
    b.ex, b.ey, b.ez = ax.hi - ax.lo, ay.hi - ay.lo, az.hi - az.lo
    return b


def face_columns(b: SwBand):
    """Number of columns of each face and their start offsets."""
    ex2, ey2 = b.ex - 2 * DEPTH, b.ey - 2 * DEPTH
    n = [b.ey * b.ez, b.ey * b.ez, ex2 * b.ez, ex2 * b.ez, ex2 * ey2, ex2 * ey2]
    return n, np.concatenate([[0], np.cumsum(n)])


class BandNet:
    """Parameters of the six face networks, one device tensor [6, P]."""

    def __init__(self, device=None, seed=0):
        self.P = param_count()
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.lay = layout()
        self.params = torch.zeros((6, self.P), dtype=torch.float32, device=self.device)
        rng = np.random.default_rng(seed)
        host = np.zeros((6, self.P), np.float32)
        for f in range(6):
            for name, (fan_in, fan_out) in (("W1", (NFEAT, NH)), ("W2", (NH, NH)), ("W3", (NH, NOUT))):
                a, b = self.lay[name]
                lim = np.sqrt(6.0 / (fan_in + fan_out))
                host[f, a:b] = rng.uniform(-lim, lim, b - a)
            a, b = self.lay["W3"]
            host[f, a:b] *= 0.1  # start close to "no correction"
            for name in ("in_scale", "out_scale"):
                a, b = self.lay[name]
                host[f, a:b] = 1.0
        self.params.copy_(torch.from_numpy(host))

    def set_scales(self, face, in_scale, out_scale):
        a, b = self.lay["in_scale"]
        self.params[face, a:b] = torch.as_tensor(in_scale, dtype=torch.float32, device=self.device)
        a, b = self.lay["out_scale"]
        self.params[face, a:b] = torch.as_tensor(out_scale, dtype=torch.float32, device=self.device)

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy())

    @classmethod
    def load(cls, path, device=None):
        net = cls(device)
        with np.load(path, allow_pickle=False) as z:
            net.params.copy_(torch.from_numpy(z["params"]))
        return net


# ------------------------------------------------------------ data generation
def random_layered_model(shape, rng, vp_range=(1800.0, 4000.0), nlayers=4, pert=0.05):
    """Random layered model along the last axis with a smooth random
    perturbation (clipped to +-2 standard deviations); vs = vp/1.73, Gardner
    density."""
    from scipy import ndimage
    nz = shape[-1]
    cuts = np.sort(rng.integers(1, nz - 1, nlayers - 1))
    vals = np.sort(rng.uniform(*vp_range, nlayers))
    prof = np.empty(nz)
    edges = np.concatenate([[0], cuts, [nz]])
    for t in range(nlayers):
        prof[edges[t]:edges[t + 1]] = vals[t]
    vp = np.broadcast_to(prof, shape).astype(np.float64)
    noise = ndimage.gaussian_filter(rng.standard_normal(shape), 3.0, mode="wrap")
    noise /= noise.std() + 1e-12
    vp = vp * (1 + pert * np.clip(noise, -2.0, 2.0))
    vs = vp / 1.73
    rho = 1741.0 * (vp / 1000.0) ** 0.25
    return vs.astype(np.float32), vp.astype(np.float32), rho.astype(np.float32)


def courant(vp_max, d, dt, order=2):
    """vp_max dt / d * sqrt(3) * sum|c|: the 3-D staggered scheme is stable below 1."""
    from .coeffs import diff_coff
    return float(vp_max) * dt / d * np.sqrt(3.0) * float(np.sum(np.abs(diff_coff(order))))


# CPML profile of every learned-boundary run, independent of the source: the
# networks learn the error of one particular thin layer.
PML_FREQ = 15.0
# Target reflection coefficient of the CPML profile (d0 = -3 vp ln R / 2L):
# the wide layers keep the 1e-6 of the reference's drivers; a layer of at most
# THIN_MAX cells uses THIN_RCOEF, since a profile that steep over 4 cells
# reflects from its own discretisation.  Reflected energy of a 4-cell layer
# relative to the 20-cell one (homogeneous 3.5 km/s model, 15 Hz moment
# source, 200 steps, CPU oracle): R = 1e-6: 8.3e4, 1e-4: 7.7e3, 1e-3: 3.2e3,
# 1e-2: 1.7e4, 1e-1: 7.0e5; kmax 2 or 4, alpha_max 0 or 30 pi at R = 1e-3
# were all worse.
WIDE_RCOEF, THIN_RCOEF, THIN_MAX = 1e-6, 1e-3, 8


def pml_rcoef(npml):
    return THIN_RCOEF if npml <= THIN_MAX else WIDE_RCOEF


def _sim(vs, vp, rho, d, dt, cell, mt, amps, npml, f, rcoef=None):
    from .Elastic3D_SSG_Tensor import ElasticWAVE
    cr = courant(vp.max(), d, dt)
    if cr >= 0.95:
        raise ValueError(f"model violates the stability limit (Courant number {cr:.3f})")
    nt = len(amps)
    mt3 = np.zeros((3, 3), np.float32)
    w = ElasticWAVE(vs, vp, rho, d, d, d, dt, *cell, np.zeros((1, 1), np.int64) + cell[0],
                    np.zeros((1, 1), np.int64) + cell[1], np.zeros((1, 1), np.int64) + cell[2], mt3, 1.0, nt, 2, f,
                    source={"kind": "moment", "cell": cell, "mt": mt, "wavelet": amps})
    p = {"vp_max": float(vp.max()), "pml_x_thick": npml * d, "pml_y_thick": npml * d, "pml_z_thick": npml * d,
         "Rcoef": pml_rcoef(npml) if rcoef is None else rcoef, "theta": 1, "alpha_max_pml": np.pi * PML_FREQ,
         "kmax_pml": 1.0}
    w.SetADEPML3D([True] * 6, p)
    return w


def grid_size(n_int, npml, s=2):
    """Grid points per axis that leave n_int interior cells with an npml-cell CPML."""
    return n_int + 2 * s + 2 * npml + 1


def collect_run(rng, n_int=40, steps=160, d=10.0, dt=1e-3, f=None, thin=4, wide=20, keep=0.2):
    """One teacher-forced run.

    A wide-CPML run of a random model is the reference; a thin-CPML run of the
    same interior is reset to the reference interior before every step.  For
    every band column: features [cur, cur - prev] (cur = thin band after the
    step, prev = the band before it) and target = reference band - cur.
    Returns per-face (X [m, 72], T [m, 36]) on the device, a random ``keep``
    fraction of the columns of every step after the first 10.  The source
    peak frequency is drawn from 10-20 Hz unless ``f`` is given."""
    lib = _sig()
    if f is None:
        f = float(rng.uniform(10.0, 20.0))
    nw, nt_ = grid_size(n_int, wide), grid_size(n_int, thin)
    off = wide - thin
    vs, vp, rho = random_layered_model((nw, nw, nw), rng)
    lo_w = 2 + wide
    src = tuple(int(v) for v in rng.integers(lo_w + 6, lo_w + n_int - 6, 3))
    m = rng.standard_normal(6)
    m /= np.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2 + 2 * (m[3] ** 2 + m[4] ** 2 + m[5] ** 2))
    amps = (1e6 * ricker2(np.arange(steps), dt, 1.0 / f, f, 1.0)).astype(np.float32)
    W = _sim(vs, vp, rho, d, dt, src, m, amps, wide, f)
    sl = slice(off, off + nt_)
    T_ = _sim(vs[sl, sl, sl].copy(), vp[sl, sl, sl].copy(), rho[sl, sl, sl].copy(), d, dt,
              tuple(c - off for c in src), m, amps, thin, f)
    bt, bw = band_of(T_._e), band_of(W._e)
    ncol = lib.sw_band_columns(C.byref(bt))
    counts, col0 = face_columns(bt)
    dev = T_._e.device
    cur = torch.empty((ncol, NOUT), device=dev)
    prv = torch.empty((ncol, NOUT), device=dev)
    wcur = torch.empty((ncol, NOUT), device=dev)
    st = T_._e.stream()
    xs = {f_: [] for f_ in range(6)}
    ts = {f_: [] for f_ in range(6)}
    names = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    # interior boxes in device layout [z][y][x]
    ti = (slice(bt.oz, bt.oz + bt.ez), slice(bt.oy, bt.oy + bt.ey), slice(bt.ox, bt.ox + bt.ex))
    wi = (slice(bw.oz, bw.oz + bw.ez), slice(bw.oy, bw.oy + bw.ey), slice(bw.ox, bw.ox + bw.ex))
    gen = torch.Generator(device=dev).manual_seed(int(rng.integers(1 << 31)))
    for n in range(steps):
        for nm in names:
            T_._e.f[nm][ti].copy_(W._e.f[nm][wi])
        L.check(lib.sw_band_gather(T_._e.descriptor(), C.byref(bt), L.fptr(prv), st), "band gather")
        T_.run(1, n)
        W.run(1, n)
        L.check(lib.sw_band_gather(T_._e.descriptor(), C.byref(bt), L.fptr(cur), st), "band gather")
        L.check(lib.sw_band_gather(W._e.descriptor(), C.byref(bw), L.fptr(wcur), st), "band gather")
        if n < 10:
            continue
        x = torch.cat([cur, cur - prv], 1)
        t = wcur - cur
        for f_ in range(6):
            a, b = int(col0[f_]), int(col0[f_ + 1])
            k = max(1, int(keep * (b - a)))
            idx = torch.randint(a, b, (k,), device=dev, generator=gen)
            xs[f_].append(x[idx])
            ts[f_].append(t[idx])
    return {f_: (torch.cat(xs[f_]), torch.cat(ts[f_])) for f_ in range(6)}


def normalise(X, T, in_scale):
    """Divide every sample by its column scale s = RMS(X / in_scale); drop
    silent samples.  Returns (X / s, T / s, s)."""
    s = torch.sqrt(torch.mean((X / in_scale) ** 2, 1))
    keep = s > 0
    X, T, s = X[keep], T[keep], s[keep]
    return X / s[:, None], T / s[:, None], s


class _Trainer:
    """HIP forward / backward / Adam of one face network (buffers for one batch size)."""

    def __init__(self, net: BandNet, face, batch):
        self.lib, self.net, self.face, self.batch = _sig(), net, face, batch
        dev = net.device
        self.ntrain = net.lay["b3"][1]  # weights and biases come first
        self.grad = torch.zeros(self.ntrain, device=dev)
        self.m1 = torch.zeros(self.ntrain, device=dev)
        self.m2 = torch.zeros(self.ntrain, device=dev)
        self.H1 = torch.empty((batch, NH), device=dev)
        self.H2 = torch.empty((batch, NH), device=dev)
        self.Y = torch.empty((batch, NOUT), device=dev)
        self.work = torch.empty(batch * (NOUT + 2 * NH), device=dev)
        self.rows = torch.empty(batch, device=dev)
        self.st = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        self.t = 0

    def step(self, xb, tb, lr):
        p = self.net.params[self.face]
        lib, B = self.lib, xb.shape[0]
        L.check(lib.sw_mlp_forward(L.fptr(p), L.fptr(xb), B, L.fptr(self.H1), L.fptr(self.H2), L.fptr(self.Y),
                                   self.st), "mlp forward")
        L.check(lib.sw_mlp_loss_backward(L.fptr(p), L.fptr(xb), L.fptr(self.H1), L.fptr(self.H2), L.fptr(self.Y),
                                         L.fptr(tb), B, L.fptr(self.grad), L.fptr(self.work), L.fptr(self.rows),
                                         self.st), "mlp backward")
        self.t += 1
        L.check(lib.sw_adam(L.fptr(p), L.fptr(self.grad), L.fptr(self.m1), L.fptr(self.m2), self.ntrain, lr, 0.9,
                            0.999, 1e-8, self.t, self.st), "adam")

    def predict(self, Xn):
        """Network output (in units of out_scale) for normalised samples, in batches."""
        p = self.net.params[self.face]
        out = torch.empty((Xn.shape[0], NOUT), device=Xn.device)
        for a in range(0, Xn.shape[0], self.batch):
            xb = Xn[a:a + self.batch].contiguous()
            B = xb.shape[0]
            L.check(self.lib.sw_mlp_forward(L.fptr(p), L.fptr(xb), B, L.fptr(self.H1), L.fptr(self.H2),
                                            L.fptr(self.Y), self.st), "mlp forward")
            out[a:a + B] = self.Y[:B]
        return out


def residual_fraction(trainer, X, T, in_scale, out_scale):
    """sum |T - correction(X)|^2 / sum |T|^2 over raw samples: the one-step
    error left after the correction, relative to the error without it."""
    s = torch.sqrt(torch.mean((X / in_scale) ** 2, 1))
    live = s > 0
    corr = torch.zeros_like(T)
    corr[live] = trainer.predict((X[live] / s[live, None]).contiguous()) * out_scale * s[live, None]
    num = torch.sum((T.double() - corr.double()) ** 2)
    return float(num / torch.sum(T.double() ** 2).clamp_min(1e-300))


def calibrate(trainer, X, T, in_scale, out_scale):
    """Least-squares output scale alpha = <T, C> / <C, C> of the correction C
    on (X, T), clamped to [0, 1]: for 0 < alpha <= 2<T,C>/<C,C> the one-step
    error |T - alpha C|^2 is below |T|^2 whenever <T, C> > 0."""
    s = torch.sqrt(torch.mean((X / in_scale) ** 2, 1))
    live = s > 0
    corr = trainer.predict((X[live] / s[live, None]).contiguous()).double() * out_scale.double() * s[live, None]
    tc = float(torch.sum(T[live].double() * corr))
    cc = float(torch.sum(corr * corr))
    return min(max(tc / cc, 0.0), 1.0) if cc > 0 else 0.0


def train_face(net: BandNet, face, X, T, Xv=None, Tv=None, iters=3000, batch=4096, lr=2e-3, seed=0, log=None,
               eval_every=250):
    """Fit one face network with the HIP training kernels.

    Samples are drawn with probability proportional to s^2 (their squared
    column scale), so the mean loss over a batch of normalised samples
    estimates the raw squared error.  With a validation set (Xv, Tv) the
    parameters with the lowest validation residual are kept, then the output
    scale is calibrated on the validation set (``calibrate``).  Returns the
    history [(iter, train residual fraction, validation residual fraction)];
    its last entry is after calibration."""
    dev = net.device
    in_scale = torch.sqrt(torch.mean(X * X, 0)) + 1e-30
    out_scale = torch.sqrt(torch.mean(T * T, 0)) + 1e-30
    net.set_scales(face, in_scale, out_scale)
    Xn, Tn, s = normalise(X, T, in_scale)
    w = (s.double() ** 2)
    w = w / w.sum()
    tr = _Trainer(net, face, batch)
    g = torch.Generator(device=dev).manual_seed(seed)
    n_eval = min(X.shape[0], 200000)

    def evaluate(os_=out_scale):
        a = residual_fraction(tr, X[:n_eval], T[:n_eval], in_scale, os_)
        b = residual_fraction(tr, Xv, Tv, in_scale, os_) if Xv is not None else float("nan")
        return a, b

    hist = [(0,) + evaluate()]
    best = (hist[0][2] if Xv is not None else hist[0][1], net.params[face].clone())
    for it in range(1, iters + 1):
        idx = torch.multinomial(w, batch, replacement=True, generator=g)
        tr.step(Xn[idx].contiguous(), Tn[idx].contiguous(), lr * (0.1 ** (it / iters)))
        if it % eval_every == 0 or it == iters:
            hist.append((it,) + evaluate())
            score = hist[-1][2] if Xv is not None else hist[-1][1]
            if score < best[0]:
                best = (score, net.params[face].clone())
            if log:
                log(f"face {FACES[face]} iter {it} train {hist[-1][1]:.4f} val {hist[-1][2]:.4f}")
    net.params[face].copy_(best[1])
    # calibrate the output scale on held-out samples (the training ones without)
    cx, ct = (Xv, Tv) if Xv is not None else (X, T)
    alpha = calibrate(tr, cx, ct, in_scale, out_scale)
    net.set_scales(face, in_scale, out_scale * alpha)
    hist.append((iters + 1,) + evaluate(out_scale * alpha))
    if log:
        log(f"face {FACES[face]} output scale {alpha:.3f}: train {hist[-1][1]:.4f} val {hist[-1][2]:.4f}")
    return hist


def train(n_runs=16, n_int=40, steps=160, iters=3000, seed=0, log=print, n_val=None, batch=4096, lr=2e-3):
    """Generate teacher-forced samples with the plain solver and train all six
    face networks; the last ``n_val`` runs (default n_runs // 8, at least 1)
    are held out for validation."""
    rng = np.random.default_rng(seed)
    t0 = time.time()
    n_val = max(1, n_runs // 8) if n_val is None else n_val
    pools = {f: ([], []) for f in range(6)}
    for r in range(n_runs):
        data = collect_run(rng, n_int=n_int, steps=steps)
        for f in range(6):
            pools[f][0].append(data[f][0])
            pools[f][1].append(data[f][1])
        if log:
            log(f"run {r + 1}/{n_runs} collected ({time.time() - t0:.1f} s)")
    net = BandNet(seed=seed)
    hists = {}
    nt = n_runs - n_val
    for f in range(6):
        X = torch.cat(pools[f][0][:nt])
        T = torch.cat(pools[f][1][:nt])
        Xv = torch.cat(pools[f][0][nt:]) if n_val else None
        Tv = torch.cat(pools[f][1][nt:]) if n_val else None
        hists[f] = train_face(net, f, X, T, Xv, Tv, iters=iters, batch=batch, lr=lr, seed=seed + f, log=log)
    return net, hists


def linear_ceiling(n_runs=16, n_val=4, n_int=28, steps=150, seed=3):
    """Share of the one-step band error that the best linear map of the same
    72 features removes, per face: least squares T ~ X W (no bias: the wave
    equation and the CPML are linear), in f64 on teacher-forced runs.
    Returns {face: (in_sample, held_out)}: ``in_sample`` fits and evaluates
    on all ``n_runs`` runs, an upper bound on what any linear map removes
    from those samples; ``held_out`` fits on the first n_runs - n_val runs
    and evaluates on the rest (negative when the fit makes the held-out error
    larger than no correction).  removed = 1 - |T - X W|^2 / |T|^2."""
    rng = np.random.default_rng(seed)
    pools = {f: ([], []) for f in range(6)}
    for _ in range(n_runs):
        data = collect_run(rng, n_int=n_int, steps=steps)
        for f in range(6):
            pools[f][0].append(data[f][0])
            pools[f][1].append(data[f][1])

    def removed(X, T, Xe, Te):
        W = torch.linalg.lstsq(X, T).solution
        return float(1.0 - torch.sum((Te - Xe @ W) ** 2) / torch.sum(Te ** 2).clamp_min(1e-300))

    out = {}
    nt = n_runs - n_val
    for f in range(6):
        X = [x.double() for x in pools[f][0]]
        T = [t.double() for t in pools[f][1]]
        Xa, Ta = torch.cat(X), torch.cat(T)
        out[FACES[f]] = (removed(Xa, Ta, Xa, Ta),
                         removed(torch.cat(X[:nt]), torch.cat(T[:nt]), torch.cat(X[nt:]), torch.cat(T[nt:])))
    return out


# ---------------------------------------------------------------- inference
class LearnedRun:
    """Runs an ElasticWAVE (3-D) instance with the learned band correction."""

    def __init__(self, sim, net: BandNet):
        self.lib = _sig()
        self.sim = sim
        self.net = net
        self.band = band_of(sim._e)
        ncol = self.lib.sw_band_columns(C.byref(self.band))
        self.prev = torch.zeros((ncol, NOUT), device=sim._e.device)

    def reset(self):
        self.prev.zero_()

    def run(self, nsteps, it0=0):
        e = self.sim._e
        g = e.descriptor()
        L.check(self.lib.sw_ssg3d_run_learned(g, self.sim._src.s, self.sim._rec.r, C.byref(self.band),
                                              L.fptr(self.net.params), L.fptr(self.prev), it0, nsteps,
                                              e.stream()), "sw_ssg3d_run_learned")


# --------------------------------------------------------------- evaluation
def tune_output_scale(net: BandNet, alphas=(1.0, 0.5, 0.25, 0.1, 0.05), n_int=28, steps=150, seed=7):
    """Closed-loop calibration: scale every face's output by the alpha in
    ``alphas`` (or 0) that gives the least reflected energy on one held-out
    model, and keep that scale.  Returns (alpha, {alpha: energy})."""
    a, b = net.lay["out_scale"]
    base = net.params[:, a:b].clone()
    energy = {}
    for al in alphas:
        net.params[:, a:b] = base * al
        energy[al] = reflection_study(net, n_int=n_int, steps=steps, seed=seed)["learned"]
    energy[0.0] = reflection_study(net, n_int=n_int, steps=steps, seed=seed)["thin"]
    best = min(energy, key=energy.get)
    net.params[:, a:b] = base * best
    return best, energy


def reflection_study(net: BandNet, n_int=40, steps=200, seed=0, d=10.0, dt=1e-3, f=15.0, thin=4, wide=20,
                     model=None, src=None, every=5, timing_steps=0, thin_rcoef=None, runner=None):
    """Reflected energy of thin CPML (with and without the network) and of the
    wide CPML, against a padded reference whose own boundary reflections
    cannot reach the interior within ``steps``.

    ``model`` is (vs, vp, rho) over the n_int^3 interior; every run extends it
    outwards by edge replication.  Energy = sum over snapshots every ``every``
    steps and over the interior box of (v - v_ref)^2 for the three velocity
    components.  ``timing_steps`` > 0 also times the thin (alone and with
    the network) and the wide runs (ms per step, HIP events).  ``thin_rcoef``
    overrides the thin layer's target reflection coefficient; ``runner(sim)``,
    if given, replaces LearnedRun(sim, net) for the corrected thin run."""
    rng = np.random.default_rng(seed)
    if model is None:
        model = random_layered_model((n_int,) * 3, rng)
    vmax = float(np.max(model[1]))
    pad_ref = int(np.ceil(vmax * dt * steps / d / 2)) + 4
    if src is None:
        src = tuple(int(v) for v in rng.integers(6, n_int - 6, 3))
    m = rng.standard_normal(6)
    m /= np.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2 + 2 * (m[3] ** 2 + m[4] ** 2 + m[5] ** 2))
    nt_all = steps + max(timing_steps, 0)
    amps = (1e6 * ricker2(np.arange(nt_all), dt, 1.2 / f, f, 1.0)).astype(np.float32)
    runs = {}
    for name, npml, extra in (("ref", wide, pad_ref), ("wide", wide, 0), ("thin", thin, 0), ("learned", thin, 0)):
        lo = 2 + npml + extra  # first interior index on every axis
        mdl = tuple(np.ascontiguousarray(np.pad(np.asarray(a, np.float32), ((lo, lo + 1),) * 3, mode="edge"))
                    for a in model)
        rc = thin_rcoef if name in ("thin", "learned") else None
        sim = _sim(*mdl, d, dt, tuple(c + lo for c in src), m, amps, npml, f, rc)
        runs[name] = (sim, lo)
    learned = LearnedRun(runs["learned"][0], net) if runner is None else runner(runs["learned"][0])
    energy = {k: 0.0 for k in ("wide", "thin", "learned")}
    ref_norm = 0.0
    comps = ("vx", "vy", "vz")
    for n0 in range(0, steps, every):
        k = min(every, steps - n0)
        for name, (sim, lo) in runs.items():
            if name == "learned":
                learned.run(k, n0)
            else:
                sim.run(k, n0)
        rs, rlo = runs["ref"]
        box = lambda lo: (slice(lo, lo + n_int),) * 3  # noqa: E731
        for cmp in comps:
            r = rs._e.f[cmp][box(rlo)]
            ref_norm += float(torch.sum(r.double() ** 2))
            for name in energy:
                sim, lo = runs[name]
                energy[name] += float(torch.sum((sim._e.f[cmp][box(lo)].double() - r.double()) ** 2))
    out = dict(energy)
    out["reference_energy"] = ref_norm
    out["ratio_learned_to_wide"] = energy["learned"] / max(energy["wide"], 1e-300)
    out["ratio_thin_to_wide"] = energy["thin"] / max(energy["wide"], 1e-300)
    if timing_steps > 0:
        for name in ("wide", "thin", "learned"):
            sim, _ = runs[name]
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            if name == "learned":
                learned.run(timing_steps, steps)
            else:
                sim.run(timing_steps, steps)
            b.record()
            torch.cuda.synchronize()
            out[f"ms_per_step_{name}"] = a.elapsed_time(b) / timing_steps
    return out


# ------------------------------------------------------------ history study
def _face_cells(b: SwBand, dev):
    """Device-layout (z, y, x) index tensors [ncol, DEPTH] of the band cells of
    every face, faces split as sw_band_gather splits them (x faces over the
    whole interior face, y faces without the x bands, z faces without both)."""
    lo, hi = (b.ox, b.oy, b.oz), (b.ox + b.ex, b.oy + b.ey, b.oz + b.ez)
    full = [torch.arange(lo[a], hi[a], device=dev) for a in range(3)]
    inner = [torch.arange(lo[a] + DEPTH, hi[a] - DEPTH, device=dev) for a in range(3)]
    depth = lambda a, side: (torch.arange(lo[a], lo[a] + DEPTH, device=dev) if side == 0  # noqa: E731
                             else torch.arange(hi[a] - DEPTH, hi[a], device=dev))
    cells = []
    for a in (0, 1, 2):  # x, y, z
        other = [q for q in (0, 1, 2) if q != a]
        rng_ = [full[q] if (a == 0 or (a == 1 and q == 2)) else inner[q] for q in other]
        A, B = torch.meshgrid(rng_[0], rng_[1], indexing="ij")
        A, B = A.reshape(-1, 1), B.reshape(-1, 1)
        for side in (0, 1):
            dd = depth(a, side)[None, :]
            xyz = [None, None, None]
            xyz[a] = dd.expand(A.shape[0], -1)
            xyz[other[0]], xyz[other[1]] = A.expand(-1, DEPTH), B.expand(-1, DEPTH)
            cells.append((xyz[2], xyz[1], xyz[0]))  # device layout [z][y][x]
    return cells


_FIELDS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


def _band_values(engine, idx):
    return torch.cat([engine.f[nm][idx] for nm in _FIELDS], 1)


class LinearHistoryRun:
    """Thin-CPML run whose band, after every step, gets X @ W added, X = [the
    band after the step, the corrected bands before the last ``hist``
    steps] (36 * (hist + 1) values per column), W one matrix per face."""

    def __init__(self, sim, maps, hist):
        self.sim, self.hist = sim, hist
        self.cells = _face_cells(band_of(sim._e), sim._e.device)
        self.maps = [m.float() for m in maps]
        self.past = [[] for _ in self.cells]

    def run(self, nsteps, it0=0):
        e = self.sim._e
        for n in range(it0, it0 + nsteps):
            pre = [_band_values(e, c) for c in self.cells]
            self.sim.run(1, n)
            for f_, c in enumerate(self.cells):
                self.past[f_] = (self.past[f_] + [pre[f_]])[-self.hist:]
                if len(self.past[f_]) < self.hist:
                    continue
                X = torch.cat([_band_values(e, c)] + self.past[f_][::-1], 1)
                corr = X @ self.maps[f_]
                for q, nm in enumerate(_FIELDS):
                    e.f[nm][c] += corr[:, q * DEPTH:(q + 1) * DEPTH]


def history_study(n_runs=12, hist=4, n_val=2, ridges=(1e-8, 1e-5, 1e-3), alphas=(1.0, 0.5), seed=11,
                  n_int=28, steps=150, test_seed=101, d=10.0, dt=1e-3, keep=0.2):
    """What the one-step features lack, measured: with the corrected bands of
    the last ``hist`` steps added to the band after the step, the one-step
    band error of teacher-forced runs becomes almost a linear function of the
    features (least squares per face, f64, held out on the last ``n_val``
    runs), but the least-squares correction is not stable in closed loop.
    Returns {"held_out": {face: removed fraction}, "closed_loop": {"ridge=r
    alpha=a": reflected energy with the correction / thin layer's alone} on
    the unseen model ``test_seed``} (removed = 1 - |T - XW|^2 / |T|^2; ridge
    relative to the mean diagonal of the column-scaled normal matrix)."""
    rng = np.random.default_rng(seed)
    pools = [[] for _ in range(6)]
    for _ in range(n_runs):
        f = float(rng.uniform(10.0, 20.0))
        nw, nt_ = grid_size(n_int, 20), grid_size(n_int, 4)
        off = 16
        vs, vp, rho = random_layered_model((nw, nw, nw), rng)
        src = tuple(int(v) for v in rng.integers(22 + 6, 22 + n_int - 6, 3))
        m = rng.standard_normal(6)
        m /= np.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2 + 2 * (m[3] ** 2 + m[4] ** 2 + m[5] ** 2))
        amps = (1e6 * ricker2(np.arange(steps), dt, 1.0 / f, f, 1.0)).astype(np.float32)
        W = _sim(vs, vp, rho, d, dt, src, m, amps, 20, f)
        sl = slice(off, off + nt_)
        T = _sim(vs[sl, sl, sl].copy(), vp[sl, sl, sl].copy(), rho[sl, sl, sl].copy(), d, dt,
                 tuple(c - off for c in src), m, amps, 4, f)
        bt, bw = band_of(T._e), band_of(W._e)
        ct, cw = _face_cells(bt, T._e.device), _face_cells(bw, W._e.device)
        ti = (slice(bt.oz, bt.oz + bt.ez), slice(bt.oy, bt.oy + bt.ey), slice(bt.ox, bt.ox + bt.ex))
        wi = (slice(bw.oz, bw.oz + bw.ez), slice(bw.oy, bw.oy + bw.ey), slice(bw.ox, bw.ox + bw.ex))
        gen = torch.Generator(device=T._e.device).manual_seed(int(rng.integers(1 << 31)))
        past = [[] for _ in range(6)]
        samples = [[] for _ in range(6)]
        for n in range(steps):
            for nm in _FIELDS:
                T._e.f[nm][ti].copy_(W._e.f[nm][wi])
            pre = [_band_values(T._e, c) for c in ct]
            T.run(1, n)
            W.run(1, n)
            for f_ in range(6):
                past[f_] = (past[f_] + [pre[f_]])[-hist:]
                if n < 10 or len(past[f_]) < hist:
                    continue
                cur = _band_values(T._e, ct[f_])
                X = torch.cat([cur] + past[f_][::-1], 1)
                k = max(1, int(keep * X.shape[0]))
                idx = torch.randint(0, X.shape[0], (k,), device=X.device, generator=gen)
                samples[f_].append((X[idx].double(), (_band_values(W._e, cw[f_]) - cur)[idx].double()))
        for f_ in range(6):
            pools[f_].append((torch.cat([a for a, _ in samples[f_]]), torch.cat([b for _, b in samples[f_]])))

    def fit(X, T, ridge):
        sc = X.abs().amax(0).clamp_min(1e-300)
        Xs = X / sc
        A = Xs.T @ Xs
        A = A + ridge * torch.trace(A) / A.shape[0] * torch.eye(A.shape[0], dtype=A.dtype, device=A.device)
        return torch.linalg.solve(A, Xs.T @ T) / sc[:, None]

    held = {}
    for f_ in range(6):
        Xt = torch.cat([p[0] for p in pools[f_][:-n_val]])
        Tt = torch.cat([p[1] for p in pools[f_][:-n_val]])
        Xv = torch.cat([p[0] for p in pools[f_][-n_val:]])
        Tv = torch.cat([p[1] for p in pools[f_][-n_val:]])
        Wt = fit(Xt, Tt, min(ridges))
        held[FACES[f_]] = float(1 - torch.sum((Tv - Xv @ Wt) ** 2) / torch.sum(Tv ** 2))
    loop = {}
    for ridge in ridges:
        maps = [fit(torch.cat([p[0] for p in pools[f_]]), torch.cat([p[1] for p in pools[f_]]), ridge)
                for f_ in range(6)]
        for alpha in alphas:
            r = reflection_study(None, n_int=n_int, steps=steps, seed=test_seed,
                                 runner=lambda sim: LinearHistoryRun(sim, [alpha * w for w in maps], hist))
            loop[f"ridge={ridge:g} alpha={alpha:g}"] = r["learned"] / r["thin"]
    return {"held_out": held, "closed_loop": loop}
