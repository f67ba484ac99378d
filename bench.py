"""Benchmark of the MI355X BES solver core (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line.  A "step" is one full pseudo-time step of the solver
over the whole mesh (all RK stages, or one backward-Euler step including its
linear solve); value = cells x steps / wall time of the timed region, with
the mesh and state resident in HBM before timing starts.  Per-kernel device
times come from HIP events recorded on the solver's own stream during the
timed region; the roofline entry uses the dominant kernel's algorithmic bytes
(every array it must touch counted once) over its average launch time.
The CPU baseline is the oracle (a restatement of the reference loops,
OpenMP over faces/cells) timed on a bounded sample on this host.
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

METRIC = "cell-updates/s + time to 1e-8 residual drop: plain FP64 vs mixed/learned"
HBM_PEAK_GBS = 8000.0


def alg_bytes(kind, nd):
    """Algorithmic bytes per cell of one launch (each array counted once)."""
    nv, k = nd + 2, nd + 1
    s = 8
    if kind == "residual":   # k_residual: residual + stage update
        return nv * s + nv * nd * s + k * (4 + 8 + 3 * nd * s) + nv * s + 2 * s + 4 + nv * s
    if kind == "gradlim":    # k_gradlim: gradient + limiter, writes phi*grad
        return nv * s + k * (4 + 2 * nd * s + s) + 2 * s + nv * nd * s
    if kind == "timestep":
        return nv * s + nd * s + 2 * s
    raise KeyError(kind)


def explicit_workload(args):
    from bes_amd import core, meshgen as mg
    from oracle import oracle as orc

    if args.workload == "explicit-bump-40k":
        md = mg.bump_channel(200, 100)
        mach = (0.5, 0.0)
        cfg = {"workload": "explicit-bump-40k", "mesh": "2D channel, 10% circular-arc bump, 200x100 quads "
               "split to 40,000 triangles", "mach": 0.5, "scheme": "RK4 (4-stage), Roe, Green-Gauss, "
               "Venkatakrishnan K=5, local dt, CFL 0.8"}
    else:
        n = args.box
        md = mg.box_tets(n[0], n[1], n[2], seed=20261202, jitter=0.1, slip_faces=("zmin",))
        mach = (0.84 * np.cos(np.radians(3.06)), 0.0, 0.84 * np.sin(np.radians(3.06)))
        cfg = {"workload": "explicit-tetbox", "mesh": f"jittered box, {n[0]}x{n[1]}x{n[2]} hexes x 6 = "
               f"{md.ncell} tetrahedra", "mach": 0.84, "alpha_deg": 3.06,
               "scheme": "RK4 (4-stage), Roe, Green-Gauss, Venkatakrishnan K=5, local dt, CFL 0.8"}
    nd = md.ndim
    finf = orc.freestream_primitive(mach, nd)
    U0 = np.tile(orc.conserved_from_primitive(finf, nd), (md.ncell, 1))
    U0 *= 1.0 + 0.01 * np.random.default_rng(3).uniform(-1, 1, U0.shape)
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U0)
        c.explicit_run(args.warmup, 0.8)
        c.synchronize()
        c.set_profiling(True)
        t0 = time.perf_counter()
        linf = c.explicit_run(args.steps, 0.8)
        c.synchronize()
        wall = time.perf_counter() - t0
        prof = c.profile()
        c.set_profiling(False)
        # same run without events, for the headline number
        c.set_state(U0)
        c.explicit_run(args.warmup, 0.8)
        c.synchronize()
        t0 = time.perf_counter()
        c.explicit_run(args.steps, 0.8)
        c.synchronize()
        wall_noev = time.perf_counter() - t0
    cells = md.ncell
    ms_step = wall_noev * 1e3 / args.steps
    value = cells * args.steps / wall_noev
    kern = {}
    for kname, (tot, cnt) in prof.items():
        avg = tot / cnt
        b = alg_bytes(kname, nd) * cells
        kern[kname] = {"avg_ms": avg, "launches": cnt, "alg_bytes": b, "GBps": b / (avg * 1e-3) / 1e9}
    dom = max(kern, key=lambda k: kern[k]["avg_ms"] * kern[k]["launches"])
    roof = {"bound": "hbm", "kernel": "k_" + dom, "achieved": round(kern[dom]["GBps"], 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(kern[dom]["GBps"] / HBM_PEAK_GBS, 4),
            "traffic": None, "alg_bytes_per_cell": alg_bytes(dom, nd),
            "kernels": {k: {"avg_us": round(v["avg_ms"] * 1e3, 2), "GBps": round(v["GBps"], 1),
                            "launches": v["launches"]} for k, v in kern.items()}}
    cfg["cells"] = cells
    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": cfg, "roofline": roof,
           "residual_linf_last": [float(x) for x in linf[-1]],
           "wall_with_events_s": wall}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_explicit(args, nd)
    return out


def cpu_baseline_explicit(args, nd):
    """The oracle's explicit step (reference loop structure, OpenMP) on a sample."""
    from bes_amd import meshgen as mg
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if nd == 3:
        md = mg.box_tets(24, 24, 20, seed=20261202, jitter=0.1, slip_faces=("zmin",))
        mach = (0.84 * np.cos(np.radians(3.06)), 0.0, 0.84 * np.sin(np.radians(3.06)))
    else:
        md = mg.bump_channel(200, 100)
        mach = (0.5, 0.0)
    finf = orc.freestream_primitive(mach, nd)
    U = np.tile(orc.conserved_from_primitive(finf, nd), (md.ncell, 1))
    m = orc.Mesh(md)
    m.explicit(U, 1, 0.8, finf, nthreads=threads)
    t0 = time.perf_counter()
    iters = 0
    while True:
        U, _ = m.explicit(U, 2, 0.8, finf, nthreads=threads)
        iters += 2
        if time.perf_counter() - t0 > args.cpu_seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": md.ncell * iters / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{iters} RK4 iterations on {md.ncell} cells ({'tetrahedra' if nd == 3 else 'triangles'}), "
                      f"{dt:.1f} s, oracle OpenMP face loop with atomics as the reference"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="explicit-tetbox",
                    choices=["explicit-tetbox", "explicit-bump-40k"])
    ap.add_argument("--box", type=int, nargs=3, default=[80, 64, 65])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is designed "
                          "(see DESIGN.md) but not built", "n_gpus": args.gpus}))
        sys.exit(2)
    out = explicit_workload(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
