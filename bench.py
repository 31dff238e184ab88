"""Benchmark: 3-D staggered-grid CPML time stepping on one MI355X.

A "step" is one full reference time step (moment-tensor injection, velocity
sweep, stress sweep, receiver gather) of the Elastic3D_SSG_Tensor class on a
synthetic model that is resident in HBM before the timed region starts.

Default workload (tensor_200): 200^3 cells, 10 m spacing, homogeneous vp=3500,
vs=2000, rho=2200, dt=1 ms, 20-cell CPML on six faces, seeded unit-norm moment
tensor (10 Hz Ricker) at the centre, 50x50 receivers at k = s+5; 1500 steps.

Prints one JSON line: Gcells/s (nx*ny*nz*steps/t), the roofline of the
dominant kernel (algorithmic HBM bytes per launch / measured launch time
against 8 TB/s; against the library's streaming probe with the sweeps' array
mixes and non-temporal accesses, the practical ceiling of a streaming kernel;
and against the read + write rate of a 1 GiB device copy) and a CPU baseline
(the oracle's C restatement with OpenMP on this host's cores, and on one
thread, on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
MFMA_F32_PEAK_TFLOPS = 157.3  # MI355X dense f32 matrix-core peak


def pml_params(vp_max, vs_max, npml, d, f, kmax=None):
    return {"vp_max": vp_max, "pml_x_thick": npml * d[0], "pml_y_thick": npml * d[1],
            "pml_z_thick": npml * d[2], "Rcoef": 1e-6, "theta": 1, "alpha_max_pml": np.pi * f,
            "kmax_pml": kmax if kmax is not None else vs_max / (5 * d[0] * f)}


def unit_mt(seed):
    m = np.random.default_rng(seed).standard_normal(6)
    m /= np.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2 + 2 * (m[3] ** 2 + m[4] ** 2 + m[5] ** 2))
    return m.astype(np.float32)


def workload(name, nt, seed=0):
    """Model, geometry and PML of a named workload (host arrays)."""
    surface = [True] * 6
    if name in ("tensor_200", "spml_200"):
        n = 200
        shape = (n, n, n)
        vs = np.full(shape, 2000.0, np.float32)
        vp = np.full(shape, 3500.0, np.float32)
        rho = np.full(shape, 2200.0, np.float32)
        d, dt, f, npml = (10.0, 10.0, 10.0), 1e-3, 10.0, 20
        src = (100, 100, 100)
        lin = np.round(np.linspace(3, n - 4, 50)).astype(np.int64)
        rx, ry = (a.ravel() for a in np.meshgrid(lin, lin, indexing="ij"))
        pml = pml_params(3500.0, 2000.0, npml, d, f, kmax=4.0)
    elif name in ("layered_600x600x400", "shard_600x600x54"):
        from seiswave_taichi_amd.models import layered_folded_model
        shape = (600, 600, 400)
        d, f, npml = (12.5, 12.5, 12.5), 10.0, 20
        vs, vp, rho = layered_folded_model(shape, d[0], seed=seed)
        vp_max = float(vp.max())
        dt = 0.45 * d[0] / (vp_max * np.sqrt(3.0))
        src = (300, 300, 200)
        lin = np.round(np.linspace(3, shape[0] - 4, 200)).astype(np.int64)
        rx, ry = (a.ravel() for a in np.meshgrid(lin, lin, indexing="ij"))
        pml = pml_params(vp_max, float(vs.max()), npml, d, f)
        if name == "shard_600x600x54":
            # one GPU's share of the 600x600x400 grid split 8 ways along z: a
            # middle slab of 50 planes plus the 2-plane halo on each side, CPML
            # on the four x and y faces only (its z faces are halo exchanges)
            k0 = 175 - 2
            vs, vp, rho = (np.ascontiguousarray(a[:, :, k0:k0 + 54]) for a in (vs, vp, rho))
            shape = (600, 600, 54)
            src = (300, 300, 27)
            surface = [True, True, True, True, False, False]
    else:
        raise ValueError(name)
    rz = np.full_like(rx, 2 + 5)
    return dict(shape=shape, vs=vs, vp=vp, rho=rho, d=d, dt=dt, f=f, src=src, rec=(rx, ry, rz), pml=pml,
                mt=unit_mt(seed), nt=nt, surface=surface)


def algorithmic_bytes(shape, s, slabs, spml=False):
    """Compulsory HBM bytes of one step (two-sweep schedule, f32).

    Per updated cell: velocity sweep 13 words (6 stresses, 3 velocities read,
    3 written, rho), stress sweep 17 words (3 velocities, 6 stresses read, 6
    written, lambda, mu).  Per (updated cell, PML direction) pair: 3 + 3 memory
    variables read and written (48 B; split-field PML 64 B).  N_u counts
    prod(n_d - 2s - 1) cells; S_u the updated cells whose d index is in the
    d slab, summed over d.
    """
    core = [n - 2 * s - 1 for n in shape]
    n_u = int(np.prod(core))
    s_u = 0
    for d, (lo, hi) in enumerate(slabs):
        n = shape[d]
        w = len([i for i in range(s, n - s - 1) if i < lo or i >= hi])
        s_u += w * n_u // core[d]
    per_pair = 64 if spml else 48
    vel = 52 * n_u + per_pair // 2 * s_u
    stress = 68 * n_u + per_pair // 2 * s_u
    return n_u, s_u, vel, stress


def split_component_bytes(e, n_u, s=2):
    """Compulsory bytes per sweep of the split-field class: the plain count
    (52 / 68 B per updated cell) plus, for every component and cell where the
    component's in-PML test holds (the OR over axes of the flag at its own
    stagger), its split values read and written instead of its old value:
    +20 B for a velocity or normal-stress component (3 splits), +12 B for a
    shear stress (2 splits).  Cells are counted over the updated core
    [s, n - s - 1) of every axis, as N_u is."""
    flags = []
    for ax, n in zip(e.pml, e.gridsize):
        core = slice(s, n - s - 1)
        flags.append((np.asarray(ax.inside, bool)[core], np.asarray(ax.inside_h, bool)[core]))

    def count(hx, hy, hz):  # cells where the x, y or z flag (integer 0 / half 1) holds
        fx, fy, fz = flags[0][hx], flags[1][hy], flags[2][hz]
        return int(fx.size * fy.size * fz.size - (~fx).sum() * (~fy).sum() * (~fz).sum())

    vel = 52 * n_u + 20 * (count(1, 0, 0) + count(0, 1, 0) + count(0, 0, 1))
    stress = 68 * n_u + 20 * 3 * count(0, 0, 0) + 12 * (count(1, 1, 0) + count(1, 0, 1) + count(0, 1, 1))
    # cells with some flag set: the shell, whose cells move the split mixes
    flags = [(fi | fh, None) for fi, fh in flags]
    shell = count(0, 0, 0)
    return {"velocity": vel, "stress": stress, "shell_cells": shell}


def blended_ceiling(streams, b_split, n_u, dom):
    """Streaming rate of the split-field sweep's mix: its plain cells at the
    plain mix's probe rate, its shell cells (all split values read and
    written) at the split mix's, weighted by their bytes."""
    shell = b_split["shell_cells"]
    per = {"velocity": (52, 112, "velocity_10r_3w", "split_velocity_16r_12w"),
           "stress": (68, 164, "stress_11r_6w", "split_stress_20r_21w")}[dom]
    b_plain, b_shell = per[0] * (n_u - shell), per[1] * shell
    t = b_plain / streams[per[2]]["GBs"] + b_shell / streams[per[3]]["GBs"]
    return round((b_plain + b_shell) / t, 1)


def traffic_from_profile(kernel_tag, workload_name):
    """HBM bytes per launch measured with rocprofv3 --pmc (committed summary), or None."""
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel_tag}_{workload_name}.json")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        return json.load(fh).get("bytes_per_launch")


def copy_calibration(gib):
    """Three device copies of a ``gib`` GiB buffer: a kernel of known traffic
    (reads and writes ``gib`` GiB each) by which rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters are scaled to bytes.  Returns the buffer's bytes."""
    import torch
    n = int(gib * (1 << 30)) // 4
    src_buf = torch.zeros(n, dtype=torch.float32, device="cuda")
    dst_buf = torch.empty_like(src_buf)
    for _ in range(3):
        dst_buf.copy_(src_buf)
    torch.cuda.synchronize()
    del src_buf, dst_buf
    return n * 4


def copy_ceiling_gbs(gib=1.0, reps=5):
    """Read + write rate of a device-to-device copy of `gib` GiB (bytes moved /
    time, HIP events on the copy's stream): the rate a pure streaming kernel
    reaches on this GPU, as a practical ceiling beside the 8 TB/s peak."""
    import torch
    n = int(gib * (1 << 30)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    t = a.elapsed_time(b) / 1e3 / reps
    del src, dst
    return 2.0 * n * 4 / t / 1e9


STREAM_MIXES = (("velocity_10r_3w", 10, 3), ("stress_11r_6w", 11, 6))
SPLIT_SHELL_MIXES = (("split_velocity_16r_12w", 16, 12), ("split_stress_20r_21w", 20, 21))


def stream_ceiling_gbs(n, reps=20, mixes=STREAM_MIXES):
    """Rates of the library's streaming probe (sw_stream_probe: every element
    of each array read or written once, 16-byte accesses) with the sweeps'
    array mixes on arrays of n floats, the practical HBM ceiling of a sweep.
    Several grid sizes are timed (HIP events on the launch stream); the best
    counts.  Returns {mix: {"GBs": best, "by_blocks": {...}}}."""
    import ctypes as C

    import torch

    from seiswave_taichi_amd import _lib as L
    lib = L.load()
    n = (n // 4) * 4
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for name, nin, nout in mixes:
        bufs = [torch.ones(n, dtype=torch.float32, device="cuda") for _ in range(nin + nout)]
        pin = (L._fp * nin)(*[L.fptr(b) for b in bufs[:nin]])
        pout = (L._fp * nout)(*[L.fptr(b) for b in bufs[nin:]])
        rates = {}
        for nblock in (1024, 2048, 4096, 8192, (n // 4 + 255) // 256):
            L.check(lib.sw_stream_probe(pin, nin, pout, nout, n, nblock, stream), "stream probe")
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(reps):
                L.check(lib.sw_stream_probe(pin, nin, pout, nout, n, nblock, stream), "stream probe")
            b.record()
            torch.cuda.synchronize()
            t = a.elapsed_time(b) / 1e3 / reps
            rates[nblock] = round((nin + nout) * n * 4 / t / 1e9, 1)
        out[name] = {"GBs": max(rates.values()), "by_blocks": rates}
        del bufs
    return out


def mfma_ceiling_tflops(steps=2048, reps=3):
    """Rate of the library's matrix-core probe (sw_mfma_probe: chains of f32
    16x16x4 MFMA products, four independent accumulators per wave, nothing
    else) at several grid sizes, the practical ceiling of the learned
    boundary's correction kernel.  Returns {"TFLOPs": best, "by_blocks": {...}}."""
    import ctypes as C

    import torch

    from seiswave_taichi_amd import _lib as L
    lib = L.load()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rates = {}
    for nblock in (1024, 2048, 4096, 8192):
        out = torch.empty(nblock * 256, dtype=torch.float32, device="cuda")
        L.check(lib.sw_mfma_probe(L.fptr(out), nblock, steps, stream), "matrix-core probe")
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            L.check(lib.sw_mfma_probe(L.fptr(out), nblock, steps, stream), "matrix-core probe")
        b.record()
        torch.cuda.synchronize()
        t = a.elapsed_time(b) / 1e3 / reps
        flops = nblock * 4 * steps * 4 * 2048.0  # waves x rounds x products x flop
        rates[nblock] = round(flops / t / 1e12, 1)
        del out
    return {"TFLOPs": max(rates.values()), "by_blocks": rates}

def cpu_baseline(w, budget_s=15.0):
    """Oracle (C restatement, OpenMP) on a bounded sample of the same workload."""
    from oracle import oracle as O
    aff = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    cores = min(aff, int(env)) if env and env.isdigit() else aff
    O.set_threads(cores)
    rx, ry, rz = w["rec"]
    o = O.Oracle3D("tensor", w["vs"], w["vp"], w["rho"], *w["d"], w["dt"], w["src"],
                   np.stack([rx, ry, rz], 1), 1600, 2, freq=w["f"], mt=w["mt"])
    o.set_pml(w["surface"], w["pml"])
    o.run(2)  # warm-up
    t0 = time.perf_counter()
    o.step(2)
    t1 = time.perf_counter() - t0
    nsteps = int(max(2, min(1400, budget_s / max(t1, 1e-3))))  # the one-thread steps follow, below 1,600
    t0 = time.perf_counter()
    o.run(nsteps, it0=3)
    el = time.perf_counter() - t0
    cells = int(np.prod(w["shape"]))
    # the same restatement on one thread, a few more steps (~5 s)
    O.set_threads(1)
    t0 = time.perf_counter()
    o.step(3 + nsteps)
    t1 = time.perf_counter() - t0
    n1 = int(max(1, min(100, 5.0 / max(t1, 1e-3))))
    t0 = time.perf_counter()
    o.run(n1, it0=4 + nsteps)
    el1 = time.perf_counter() - t0
    O.set_threads(cores)
    return {"value": round(cells * nsteps / el / 1e9, 4), "unit": "Gcells/s", "cores": cores,
            "kind": "port",
            "sample": f"{nsteps} steps of the same workload after 3 warm-up steps, f32, "
                      f"{cores} OpenMP threads ({el:.1f} s)",
            "one_thread": {"value": round(cells * n1 / el1 / 1e9, 5), "unit": "Gcells/s", "cores": 1,
                           "sample": f"{n1} further steps on one thread ({el1:.1f} s)"}}


def gradient_workload(args):
    """Misfit gradient of the 2-D adjoint at full size: 800 x 300 cells, 10 m,
    layered background with two +-10 % Gaussian vp/vs anomalies (150 m), 16
    explosive 10 Hz Ricker shots and 400 receivers (vx, vz) just below the top
    CPML, 2500 steps, checkpoint every 50 steps, L2 misfit against the data of
    the anomalous model.  Times one forward+backward (seconds per gradient)
    and checks a few cells against central differences of the same f32
    forward (step 1e-2 relative)."""
    import torch

    from seiswave_taichi_amd.adjoint import Elastic2DModeling, seismograms
    from seiswave_taichi_amd.models import cfl_dt, layered_2d_with_anomalies

    torch.cuda.set_device(0)
    nx, nz, d, nt, npml, f = 800, 300, 10.0, 2500, 20, 10.0
    (vp0, vs0, rho), (vp1, vs1, _) = layered_2d_with_anomalies((nx, nz), d, seed=0)
    vmax = float(vp1.max())
    dt = cfl_dt(vmax, (d, d))
    pml = {"vp_max": vmax, "pml_x_thick": npml * d, "pml_z_thick": npml * d, "Rcoef": 1e-6, "theta": 1,
           "alpha_max_pml": np.pi * f, "kmax_pml": 1.0}
    top = 2 + npml + 2
    shots = [(int(x), top) for x in np.round(np.linspace(40, nx - 41, 16))]
    rec = np.stack([np.round(np.linspace(30, nx - 31, 400)).astype(np.int64), np.full(400, top)], 1)
    # unit-scale seismograms: a misfit of 1e-9 velocities would push the f32
    # adjoint into the denormal range
    from seiswave_taichi_amd.coeffs import ricker2
    wav = 1e8 * ricker2(np.arange(nt), dt, 1.5 / f, f, 1.0)
    m = Elastic2DModeling((nx, nz), d, d, dt, nt, 2, pml, shots, rec, freq=f, checkpoint=50, wavelet=wav)
    dobs, _ = m.model(torch.from_numpy(vp1), torch.from_numpy(vs1), torch.from_numpy(rho))
    dobs = dobs.clone()
    tv = [torch.from_numpy(a).cuda().requires_grad_(True) for a in (vp0, vs0, rho)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    data = seismograms(*tv, m)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    loss = 0.5 * torch.sum((data.double() - dobs.double()) ** 2)
    loss.backward()
    torch.cuda.synchronize()
    t_all = time.perf_counter() - t0
    grads = [t.grad.double().cpu().numpy() for t in tv]
    if not all(np.all(np.isfinite(g)) for g in grads):
        raise RuntimeError("non-finite gradient")

    def misfit(vp, vs, rh):
        dd, _ = m.model(torch.from_numpy(vp), torch.from_numpy(vs), torch.from_numpy(rh))
        return 0.5 * float(torch.sum((dd.double() - dobs.double()) ** 2))

    rng = np.random.default_rng(1)
    lo, hi = 2 + npml + 8, nz - 2 - npml - 8
    fd_rel = {}
    for k, name in enumerate(("vp", "vs", "rho")):
        base = [vp0, vs0, rho]
        fd, ad = [], []
        for _ in range(args.fd_cells):
            ci, cj = int(rng.integers(2 + npml + 8, nx - 2 - npml - 8)), int(rng.integers(lo, hi))
            h = 1e-2 * float(base[k][ci, cj])
            plus = [a.copy() for a in base]
            minus = [a.copy() for a in base]
            plus[k][ci, cj] += h
            minus[k][ci, cj] -= h
            fd.append((misfit(*plus) - misfit(*minus)) / (2 * h))
            ad.append(grads[k][ci, cj])
        fd, ad = np.array(fd), np.array(ad)
        fd_rel[name] = float(np.linalg.norm(ad - fd) / max(np.linalg.norm(fd), 1e-300))
    cells = nx * nz * nt * len(shots)
    out = {"metric": "2D adjoint gradient seconds", "value": round(t_all, 3), "unit": "s/gradient",
           "higher_is_better": False, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "gradient_800x300: 800x300 2-D VS, 16 explosive shots, 400 receivers (vx, vz), "
                                  "2500 steps, checkpoint 50, L2 misfit", "dt": dt},
           "forward_s": round(t_fwd, 3), "backward_s": round(t_all - t_fwd, 3),
           "forward_Gcells_per_s": round(cells / t_fwd / 1e9, 3),
           "misfit": float(loss.item()),
           "fd_check": {"cells_per_parameter": args.fd_cells, "step": "1e-2 relative, f32 forward",
                        "relative_error": fd_rel}}
    print(json.dumps(out), flush=True)


def vs2d_workload(args):
    """2-D VS scheme at 400 x 400 (10 m, vp 3000, vs 1732, rho 2000, dt 1 ms,
    accuracy 2, 20-cell CPML on four sides), explosive 15 Hz Ricker2 at
    (200, 200), 200 receivers at j = 7 recording vx and vz: time per step of a
    1000-step run (13 MB per step, cache resident: not an HBM case)."""
    import torch

    from seiswave_taichi_amd.Elastic2D_CPML import ElasticWAVE
    from seiswave_taichi_amd.coeffs import ricker2

    torch.cuda.set_device(0)
    n, f, dt = 400, 15.0, 1e-3
    nt = args.warmup + args.steps
    vs, vp, rho = (np.full((n, n), v, np.float32) for v in (1732.0, 3000.0, 2000.0))
    rx = np.round(np.linspace(3, n - 4, 200)).astype(np.int64)
    amps = ricker2(np.arange(nt), dt, 1.5 / f, f, 1.0)
    w = ElasticWAVE(vs, vp, rho, 10.0, 10.0, dt, 200, 200, 0, 0, nt, 2, f,
                    receivers=np.stack([rx, np.full_like(rx, 7)], 1), components=("vx", "vz"),
                    source={"kind": "explosive", "cell": (200, 200), "wavelet": amps})
    w.SetADEPML2D([True] * 4, {"vp_max": 3000.0, "pml_x_thick": 200.0, "pml_z_thick": 200.0, "Rcoef": 1e-6,
                               "theta": 1, "alpha_max_pml": np.pi * f, "kmax_pml": 1732.0 / (5 * 10.0 * f)})
    w.run("VS", args.warmup, 0)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    w.run("VS", args.steps, args.warmup)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b)
    if not np.all(np.isfinite(w.data.to_numpy())):
        raise RuntimeError("non-finite seismogram")
    out = {"metric": "2D VS ms/step", "value": round(ms / args.steps, 5), "unit": "ms/step",
           "higher_is_better": False, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "dtype": "f32",
           "data": "synthetic", "Gcells_per_s": round(n * n * args.steps / (ms / 1e3) / 1e9, 3),
           "config": {"workload": "vs2d_400: 400x400 2-D VS, explosive 15 Hz, 200 receivers (vx, vz), CPML20"}}
    print(json.dumps(out), flush=True)


def learned_workload(args):
    """Learned thin boundary on the 200^3 case: train the six face networks
    with the HIP training kernels on teacher-forced runs of 32 random layered
    28^3 interiors (1,500 Adam steps per face), calibrate the output scale in
    closed loop on one held-out model, then run the homogeneous 155^3 interior
    of the 200^3 case (vp 3500, vs 2000, rho 2200, 10 m, dt 1 ms, seeded unit
    moment tensor at the centre, 10 Hz) for 300 steps with the 20-cell CPML
    (200^3 grid), the thin 4-cell CPML alone and the thin CPML with the
    network (168^3 grids), each against a padded reference run: reflected
    energy = sum over snapshots every 5 steps of the squared velocity error
    in the interior box.  Time per step of the wide and learned runs over the
    next ``steps`` steps (HIP events).  ``value`` is the speedup of the
    learned run over the 20-cell CPML; the target is a speedup at reflected
    energy <= 1.5x the 20-cell CPML's."""
    import torch

    from seiswave_taichi_amd import learned_abc as LA

    torch.cuda.set_device(0)
    torch.manual_seed(0)
    t0 = time.perf_counter()
    net, hists = LA.train(n_runs=32, n_int=28, steps=150, iters=1500, seed=1, log=None)
    alpha, _ = LA.tune_output_scale(net, n_int=28, steps=150, seed=7)
    t_train = time.perf_counter() - t0
    model = tuple(np.full((155,) * 3, v, np.float32) for v in (2000.0, 3500.0, 2200.0))
    r = LA.reflection_study(net, n_int=155, steps=300, seed=0, f=10.0, model=model, src=(77, 77, 77),
                            timing_steps=args.steps)
    ceiling = LA.linear_ceiling()
    history = LA.history_study()
    # the plain CPML at about the wide layer's reflection: 16 cells, R = 1e-3 (192^3 grid)
    r16 = LA.reflection_study(net, n_int=155, steps=300, seed=0, f=10.0, model=model, src=(77, 77, 77),
                              thin=16, thin_rcoef=1e-3, timing_steps=args.steps)
    speedup = r["ms_per_step_wide"] / r["ms_per_step_learned"]
    refl_wide = r["learned"] / r["wide"]
    out = {"metric": "learned-ABC vs 20-cell CPML speedup at equal reflection", "value": round(speedup, 4),
           "unit": "x", "higher_is_better": True, "n_gpus": 1, "steps": args.steps, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "learned_200: 155^3 homogeneous interior of the 200^3 case, 20-cell CPML "
                                  "(200^3) vs 4-cell CPML + per-face MLP (168^3), 300 steps, 10 Hz moment tensor"},
           "reflection_vs_wide": round(refl_wide, 2), "reflection_vs_thin": round(r["learned"] / r["thin"], 4),
           "thin_reflection_vs_wide": round(r["ratio_thin_to_wide"], 2), "speedup_vs_wide": round(speedup, 4),
           "ms_per_step": {"wide_20_cell": round(r["ms_per_step_wide"], 5),
                           "thin_4_cell": round(r["ms_per_step_thin"], 5),
                           "thin_4_cell_plus_network": round(r["ms_per_step_learned"], 5),
                           "plain_16_cell_R1e-3": round(r16["ms_per_step_thin"], 5)},
           "plain_16_cell_cpml": {"reflection_vs_wide": round(r16["ratio_thin_to_wide"], 3),
                                  "speedup_vs_wide": round(r16["ms_per_step_wide"] / r16["ms_per_step_thin"], 4)},
           "target": {"reflection_vs_wide_max": 1.5, "speedup_min": 1.0,
                      "met": bool(refl_wide <= 1.5 and speedup > 1.0)},
           "output_scale": alpha, "train_s": round(t_train, 1),
           "linear_map_removes_of_one_step_error": {
               "in_sample_upper_bound": {k: round(v[0], 4) for k, v in ceiling.items()},
               "held_out": {k: round(v[1], 4) for k, v in ceiling.items()}},
           "held_out_one_step_residual": [round(h[-1][2], 4) for h in hists.values()],
           "linear_map_with_4_steps_of_band_history": {
               "held_out_removes_of_one_step_error": {k: round(v, 5) for k, v in history["held_out"].items()},
               "closed_loop_reflection_vs_thin": {k: float(f"{v:.4g}") for k, v in history["closed_loop"].items()}}}
    print(json.dumps(out), flush=True)


def learned_step_workload(args):
    """Cost of the learned boundary's step on the 200^3 case: the thin 4-cell
    CPML grid of its 155^3 interior (168^3) stepped with the band correction
    after every stress sweep, for ``steps`` steps after ``warmup``, against
    the same grid without the correction and the 20-cell CPML grid (200^3).
    The network is the seeded initialisation with its output scale set to
    zero: the kernel evaluates every layer of every column as always (its
    cost does not depend on the weights), and the fields stay those of the
    thin run, finite for any number of steps (an untrained network with a
    non-zero output scale drives them to NaN within ~150 steps, and a NaN
    state is no cost measurement).  The correction kernel is also timed alone: 136,838 band columns
    per launch, each a 72-64-64-36 network (11,072 multiply-adds, the last
    layer padded to 48 outputs on the matrix cores), so its roofline is the
    f32 matrix-core rate."""
    import ctypes as C

    import torch

    from seiswave_taichi_amd import _lib as L
    from seiswave_taichi_amd import learned_abc as LA
    from seiswave_taichi_amd.coeffs import ricker2

    torch.cuda.set_device(0)
    n_int, d, dt, f = 155, 10.0, 1e-3, 10.0
    nt = args.warmup + args.steps
    amps = (1e6 * ricker2(np.arange(nt), dt, 1.2 / f, f, 1.0)).astype(np.float32)
    mt = unit_mt(0)
    sims = {}
    for name, npml in (("wide_20_cell", 20), ("thin_4_cell", 4)):
        n = LA.grid_size(n_int, npml)
        mdl = [np.full((n,) * 3, v, np.float32) for v in (2000.0, 3500.0, 2200.0)]
        c = n // 2
        sims[name] = LA._sim(*mdl, d, dt, (c, c, c), mt, amps, npml, f)
    net = LA.BandNet(seed=0)
    a, b = net.lay["out_scale"]
    net.params[:, a:b] = 0.0
    run = LA.LearnedRun(sims["thin_4_cell"], net)
    ev = lambda: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))  # noqa: E731
    ms = {}
    for name in ("wide_20_cell", "thin_4_cell", "thin_4_cell_plus_network"):
        sim = sims["thin_4_cell" if name != "wide_20_cell" else name]
        sim.reset()
        run.reset()
        go = (lambda k, i0: run.run(k, i0)) if name.endswith("network") else (lambda k, i0: sim.run(k, i0))
        go(args.warmup, 0)
        t0, t1 = ev()
        t0.record()
        go(args.steps, args.warmup)
        t1.record()
        torch.cuda.synchronize()
        ms[name] = t0.elapsed_time(t1) / args.steps
    e = sims["thin_4_cell"]._e
    if not all(bool(torch.isfinite(t).all()) for t in e.f.values()):
        raise RuntimeError("non-finite fields after the learned run")
    lib, g, stream = L.load(), e.descriptor(), e.stream()
    ncol = lib.sw_band_columns(C.byref(run.band))
    for _ in range(5):
        L.check(lib.sw_band_correct(g, C.byref(run.band), L.fptr(net.params), L.fptr(run.prev), stream), "correct")
    evs = []
    for _ in range(args.kernel_reps):  # each launch between its own events; the median counts
        t0, t1 = ev()
        t0.record()
        L.check(lib.sw_band_correct(g, C.byref(run.band), L.fptr(net.params), L.fptr(run.prev), stream), "correct")
        t1.record()
        evs.append((t0, t1))
    torch.cuda.synchronize()
    k_ms = float(np.median([a.elapsed_time(b) for a, b in evs]))
    flops = 2.0 * ncol * (72 * 64 + 64 * 64 + 64 * 36)
    flops_mfma = 2.0 * ncol * (72 * 64 + 64 * 64 + 64 * 48)
    ach = flops / (k_ms / 1e3) / 1e12
    # per column: the 36 band values and the 36 of the previous step read,
    # the 36 corrected values written to the fields and kept for the next step
    mem = 4 * 4 * 36 * ncol
    traffic = traffic_from_profile("band_correct", "learned_200_step")
    out = {"metric": "learned boundary step cost", "value": round(ms["wide_20_cell"] / ms["thin_4_cell_plus_network"], 4),
           "unit": "x (speedup of the learned step over the 20-cell CPML step)", "higher_is_better": True,
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "learned_200_step: 155^3 homogeneous interior of the 200^3 case; 20-cell CPML "
                                  "(200^3) vs 4-cell CPML (168^3) with and without the band correction; "
                                  "seeded untrained network"},
           "ms_per_step": {k: round(v, 5) for k, v in ms.items()},
           "band_columns": ncol,
           "roofline": {"bound": "mfma", "kernel": "k_band_correct", "achieved": round(ach, 2),
                        "peak": MFMA_F32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / MFMA_F32_PEAK_TFLOPS, 4),
                        "flops_per_launch": flops, "mfma_flops_per_launch": flops_mfma,
                        "launch_ms": round(k_ms, 5), "traffic": traffic},
           "memory": {"algorithmic_bytes_per_launch": mem, "GBs": round(mem / (k_ms / 1e3) / 1e9, 1),
                      "traffic_GBs": round(traffic / (k_ms / 1e3) / 1e9, 1) if traffic else None,
                      "frac_of_hbm_peak": round((traffic or mem) / (k_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}}
    if args.copy_calibration > 0:
        out["copy_calibration_bytes"] = copy_calibration(args.copy_calibration)
    probe = mfma_ceiling_tflops()
    out["roofline"]["mfma_probe_TFLOPs"] = probe["TFLOPs"]
    out["roofline"]["frac_of_mfma_probe"] = round(flops_mfma / (k_ms / 1e3) / 1e12 / probe["TFLOPs"], 4)
    out["mfma_probe"] = probe
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="tensor_200",
                    choices=["tensor_200", "spml_200", "layered_600x600x400", "shard_600x600x54", "gradient_800x300",
                             "vs2d_400", "learned_200", "learned_200_step"])
    ap.add_argument("--fd-cells", type=int, default=4,
                    help="gradient_800x300: cells per parameter of the finite-difference spot check")
    ap.add_argument("--kernel-reps", type=int, default=200)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="after the run, copy a buffer of this many GiB 3 times (a known-size kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU domain decomposition is "
                                   "a design, not built", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.config == "gradient_800x300":
        return gradient_workload(args)
    if args.config == "vs2d_400":
        return vs2d_workload(args)
    if args.config == "learned_200":
        return learned_workload(args)
    if args.config == "learned_200_step":
        return learned_step_workload(args)

    import torch

    from seiswave_taichi_amd import _lib as L
    from seiswave_taichi_amd.Elastic3D_SSG_SPML import ElasticWAVE as SPML
    from seiswave_taichi_amd.Elastic3D_SSG_Tensor import ElasticWAVE as Tensor
    from seiswave_taichi_amd.coeffs import ricker2

    torch.cuda.set_device(0)
    nt = args.warmup + args.steps
    w = workload(args.config, nt)
    nx, ny, nz = w["shape"]
    rx, ry, rz = w["rec"]
    amps = ricker2(np.arange(nt), w["dt"], 1.5 / w["f"], w["f"], 1.0)
    src = {"kind": "moment", "cell": w["src"], "mt": w["mt"], "wavelet": amps}
    spml = args.config == "spml_200"
    if spml:
        sim = SPML(w["vs"], w["vp"], w["rho"], *w["d"], w["dt"], *w["src"], 0, 0, 0, nt, 2, w["f"],
                   source=src, receivers=np.stack([rx, ry, rz], 1))
    else:
        mt3 = np.zeros((3, 3), np.float32)
        m = w["mt"]
        mt3[0, 0], mt3[1, 1], mt3[2, 2], mt3[0, 1], mt3[0, 2], mt3[1, 2] = m
        sim = Tensor(w["vs"], w["vp"], w["rho"], *w["d"], w["dt"], *w["src"], rx[None], ry[None], rz[None],
                     mt3, 1.0, nt, 2, w["f"], source=src)
    sim.SetADEPML3D(w["surface"], w["pml"])
    torch.cuda.synchronize()

    sim.run(args.warmup, 0)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    sim.run(args.steps, args.warmup)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    t = ms / 1e3
    cells = nx * ny * nz
    gcells = cells * args.steps / t / 1e9
    trace = sim.data.to_numpy()
    if not np.all(np.isfinite(trace)):
        raise RuntimeError("non-finite seismogram")

    # per-kernel launch times (same stream, HIP events), for the roofline: the
    # two sweeps alternate as in a step, so each launch meets the caches the
    # other sweep leaves (a sweep repeated back to back re-reads what it just
    # wrote and runs faster than inside the stepped run)
    e = sim._e
    g = e.descriptor()
    stream = e.stream()
    lib = L.load()
    vel = lib.sw_ssg3d_velocity_spml if spml else lib.sw_ssg3d_velocity
    stre = lib.sw_ssg3d_stress_spml if spml else lib.sw_ssg3d_stress
    for _ in range(5):
        L.check(vel(g, stream), "velocity")
        L.check(stre(g, stream), "stress")
    evs = []
    for _ in range(args.kernel_reps):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ev[0].record()
        L.check(vel(g, stream), "velocity")
        ev[1].record()
        L.check(stre(g, stream), "stress")
        ev[2].record()
        evs.append(ev)
    torch.cuda.synchronize()
    kt = {"velocity": float(np.mean([v[0].elapsed_time(v[1]) for v in evs])) / 1e3,
          "stress": float(np.mean([v[1].elapsed_time(v[2]) for v in evs])) / 1e3}
    slabs = [(ax.lo, ax.hi) for ax in e.pml]
    n_u, s_u, b_vel, b_str = algorithmic_bytes(w["shape"], 2, slabs, spml)
    dom = "stress" if kt["stress"] >= kt["velocity"] else "velocity"
    b_dom = b_str if dom == "stress" else b_vel
    ach = b_dom / kt[dom] / 1e9
    tag = {"stress": "ssg3d_stress", "velocity": "ssg3d_velocity"}[dom] + ("_spml" if spml else "")
    traffic = traffic_from_profile(tag, args.config)
    step_bytes = b_vel + b_str
    out = {
        "metric": "3D SSG Gcells/s",
        "value": round(gcells, 3),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {nx}x{ny}x{nz} Elastic3D_SSG_"
                               f"{'SPML' if spml else 'Tensor'} CPML20, order 4, moment tensor, "
                               f"{len(rx)} receivers", "grid": [nx, ny, nz], "accuracy": 2,
                   "pml_cells": 20, "pml_faces": int(sum(w["surface"])), "receivers": int(len(rx))},
        "roofline": {"bound": "hbm", "kernel": tag, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": b_dom, "launch_ms": round(kt[dom] * 1e3, 4)},
        "step_hbm_frac": round(step_bytes / (ms / args.steps / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernels_ms": {k: round(v * 1e3, 4) for k, v in kt.items()},
        "bytes_per_step": step_bytes, "N_u": n_u, "S_u": s_u,
    }
    if spml:
        # The split-field class updates a component from its split values
        # wherever the component's own in-PML test holds
        # (src/Elastic3D_SSG_SPML.py:182-391), in all its split directions,
        # not only the damped ones the per-pair count assumes: there the old
        # value is not read and 3 (normal components) or 2 (shears) split
        # values are read and written, +20 B or +12 B per such cell.
        b_split = split_component_bytes(e, n_u)
        a_split = b_split[dom] / kt[dom] / 1e9
        out["roofline_split_components"] = {
            "algorithmic_bytes_per_launch": int(b_split[dom]), "shell_cells": b_split["shell_cells"],
            "algorithmic_bytes_by_sweep": {k: int(b_split[k]) for k in ("velocity", "stress")},
            "achieved": round(a_split, 1), "frac": round(a_split / HBM_PEAK_GBS, 4)}
    ceiling = copy_ceiling_gbs()
    out["roofline"]["copy_ceiling_GBs"] = round(ceiling, 1)
    out["roofline"]["frac_of_copy_ceiling"] = round(ach / ceiling, 4)
    streams = stream_ceiling_gbs(cells, mixes=STREAM_MIXES + (SPLIT_SHELL_MIXES if spml else ()))
    mix = streams["stress_11r_6w" if dom == "stress" else "velocity_10r_3w"]["GBs"]
    out["roofline"]["stream_ceiling_GBs"] = mix
    out["roofline"]["frac_of_stream_ceiling"] = round(ach / mix, 4)
    out["stream_probe"] = streams
    if spml:
        rs = out["roofline_split_components"]
        rs["stream_ceiling_GBs"] = blended_ceiling(streams, b_split, n_u, dom)
        rs["frac_of_stream_ceiling"] = round(rs["achieved"] / rs["stream_ceiling_GBs"], 4)
    if args.copy_calibration > 0:
        out["copy_calibration_bytes"] = copy_calibration(args.copy_calibration)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
