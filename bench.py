"""Benchmark: cells x BDF-steps / s of the per-timestep hot path on one GPU.

Workload (default): 3-D 256^3 periodic polycrystal, 64 Voronoi grains with
S3-uniform orientations, solid fraction 0.7, binary CALPHAD KKS at 1500 K with
seeded Redlich-Kister parameters (physical units, dx = 0.02 um).  The same initial
state is advanced by the plain FP64 path and by the mixed path (FP32 FAC
preconditioner under the FP64 Newton-Krylov outer loop); value is the mixed
path's throughput, with its per-field relative L2 error against the FP64 run
after the same steps.  dt = 1e-5 s; the warmup steps carry the initial
transient (white noise in c relaxes on ~1e-8 s, so the first steps are cut and
h regrows to dt), the timed steps run at dt.

Prints one JSON line (see the keys below).  Run with --gpus 1 only: this build
targets a single MI355X.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from ampe_amd import params as P  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(n, seed=1002):
    p = P.polycrystal_params(ndim=3, seed=seed)
    t = time.time()
    phi, c, q = polycrystal_state_fast(n, seed=seed)
    log(f"[bench] initial state {n}^3 in {time.time() - t:.1f}s")
    return p, phi, c, q


def polycrystal_state_fast(n, seed=1002, ngrains=64, solid_fraction=0.7, width=3.0):
    """Same construction as ampe_amd.params.polycrystal_state, with a periodic
    k-d tree for the nearest-seed search."""
    from scipy.spatial import cKDTree
    rng = np.random.Generator(np.random.Philox(seed))
    seeds = rng.uniform(0, n, (ngrains, 3))
    quats = P.random_unit_quaternions(rng, ngrains)
    tree = cKDTree(seeds, boxsize=float(n))
    x = np.arange(n, dtype=np.float64) + 0.5
    dist = np.empty((n, n, n))
    label = np.empty((n, n, n), dtype=np.int64)
    for k in range(n):
        zz, yy, xx = np.meshgrid([x[k]], x, x, indexing="ij")
        pts = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1)
        d, l = tree.query(pts, workers=-1)
        dist[k] = d.reshape(n, n)
        label[k] = l.reshape(n, n)
    thr = np.quantile(dist, solid_fraction)
    phi = 0.5 * (1.0 - np.tanh((dist - thr) / (width / 2.0)))
    c = rng.uniform(0.3, 0.5, (n, n, n))
    q = quats[:, label]
    return phi, c, q


def solver(prec, args):
    return P.solver_defaults(dt=args.dt, maxl=args.maxl, max_newton=5, atol=args.atol,
                             rtol=args.atol * 1e-2, precision=prec, fac_sweeps=2,
                             quat_sweeps=2)


def run_gpu(p, s, state, n, warmup, steps):
    import torch
    from ampe_amd import hip as H
    box = H.boxes_of((n, n, n), 3)
    sim = H.Sim(p, s, [box], periodic=box)
    sim.set_state(*state)
    if warmup:
        rc, _ = sim.advance(warmup)
        if rc:
            raise RuntimeError("warmup step failed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rc, st = sim.advance(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if rc:
        raise RuntimeError(f"timed steps failed: {st}")
    dev_ms = sim.last_time_ms()
    fields = sim.get_state((n, n, n))
    return dict(wall=wall, dev_ms=dev_ms, stats=st, fields=fields)


def kernel_roofline(p, state, n, reps=20):
    """Average duration of the fused phase-RHS kernel on the workload's level,
    timed with events on the stream it is launched on."""
    import torch
    from ampe_amd import hip as H
    dev = torch.device("cuda:0")
    box = H.boxes_of((n, n, n), 3)
    lev = H.Level(3, [box], 1, periodic=box)
    g = n + 2

    def ghosted(a, ncomp=1):
        out = np.zeros((ncomp, g, g, g))
        out[:, 1:-1, 1:-1, 1:-1] = np.asarray(a).reshape(ncomp, n, n, n)
        return torch.from_numpy(out).to(dev)

    phi, c, q = state
    tphi, tq = ghosted(phi), ghosted(q, 4)
    tcl, tcs = ghosted(c), ghosted(c)
    rhs = torch.zeros_like(tphi)
    stream = torch.cuda.current_stream()
    sp = C.c_void_p(stream.cuda_stream)
    ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    H.check(H.lib().ampe_ghost_fill(lev.h, ptr(tphi), 1, sp))
    H.check(H.lib().ampe_ghost_fill(lev.h, ptr(tq), 4, sp))
    for _ in range(3):
        H.check(H.lib().ampe_phase_rhs(lev.h, C.byref(p), ptr(tphi), ptr(tq), ptr(tcl),
                                       ptr(tcs), ptr(rhs), sp))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        H.check(H.lib().ampe_phase_rhs(lev.h, C.byref(p), ptr(tphi), ptr(tq), ptr(tcl),
                                       ptr(tcs), ptr(rhs), sp))
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    # algorithmic bytes per cell: read phi, q(4), cL, cS; write rhs (FP64)
    bytes_per_cell = 8 * (1 + 4 + 2) + 8
    achieved = bytes_per_cell * n ** 3 / (ms * 1e-3) / 1e9
    return dict(kernel="k_phase_rhs<3>", ms=ms, bytes_per_cell=bytes_per_cell,
                achieved=achieved)


def cpu_baseline(args):
    """The FP64 C oracle (OpenMP) on the same workload at a bounded size:
    the same untimed warmup steps, then cpu_steps timed steps."""
    from oracle import oracle as O
    n = args.cpu_n
    p, phi, c, q = make_workload(n, seed=1002)
    s = solver(P.FP64, args)
    sim = O.Sim(p, s, n)
    sim.set_state(phi, c, q)
    rc, st0 = sim.advance(args.warmup)
    t0 = time.perf_counter()
    rc, st = sim.advance(args.cpu_steps)
    dt = time.perf_counter() - t0
    steps = st["nst"] - st0["nst"]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=n ** 3 * steps / dt, unit="cells*steps/s", cores=cores,
                kind="port",
                sample=f"oracle FP64 integrator (OpenMP), same workload at {n}^3 after "
                       f"{args.warmup} warmup steps: {steps} timed BDF steps in {dt:.1f}s, "
                       f"rc={rc}")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--dt", type=float, default=1.0e-5)
    ap.add_argument("--atol", type=float, default=1.0e-6)
    ap.add_argument("--maxl", type=int, default=10)
    ap.add_argument("--cpu-n", type=int, default=32)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only (this build has no multi-GPU path); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.init()

    n = args.n
    p, phi, c, q = make_workload(n)
    state = (phi, c, q)
    log("[bench] FP64 run")
    r64 = run_gpu(p, solver(P.FP64, args), state, n, args.warmup, args.steps)
    log(f"[bench] FP64: {r64['wall']:.2f}s {r64['stats']}")
    log("[bench] mixed run")
    rmx = run_gpu(p, solver(P.FAC_FP32, args), state, n, args.warmup, args.steps)
    log(f"[bench] mixed: {rmx['wall']:.2f}s {rmx['stats']}")
    cells = n ** 3
    v64 = cells * args.steps / r64["wall"]
    vmx = cells * args.steps / rmx["wall"]
    errs = {name: rel_l2(a, b) for name, a, b in
            zip(("phi", "c", "q"), rmx["fields"][:3], r64["fields"][:3])}
    roof = kernel_roofline(p, state, n)
    log(f"[bench] roofline {roof}")
    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(args)
        log(f"[bench] cpu {cpu}")
    same_t = abs(r64["stats"]["t"] - rmx["stats"]["t"]) <= 1e-12 * max(r64["stats"]["t"], 1e-300)
    out = {
        "metric": "cells*BDF-steps/s at <=1e-6 L2 vs FP64; mixed vs plain FP64, 1 GPU",
        "value": vmx,
        "unit": "cells*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rmx["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FAC preconditioner f32)",
        "data": "synthetic",
        "config": {"workload": f"3D {n}^3 periodic CALPHAD-KKS polycrystal, 64 Voronoi "
                               f"grains, solid fraction 0.7, T=1500K, dt={args.dt}, "
                               f"atol={args.atol}, maxl={args.maxl}",
                   "cells": cells, "parallelism": "single GPU"},
        "fp64_value": v64,
        "speedup_mixed_vs_fp64": vmx / v64,
        "rel_l2_mixed_vs_fp64": errs,
        "same_step_sequence": same_t,
        "counters_fp64": r64["stats"],
        "counters_mixed": rmx["stats"],
        "roofline": {"bound": "hbm", "achieved": roof["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": roof["achieved"] / HBM_PEAK_GBS,
                     "traffic": None, "kernel": roof["kernel"],
                     "kernel_ms": roof["ms"], "bytes_per_cell": roof["bytes_per_cell"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
