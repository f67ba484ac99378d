"""Benchmark of the MI355X BES solver core (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints one JSON line on stdout (progress goes to stderr).

A "step" is one full pseudo-time step over the whole mesh: one backward-Euler
step including its Jacobian assembly and linear solve (implicit workloads), or
all RK stages (explicit workloads).  value = cells x K / wall time of exactly K
timed steps, with mesh and state resident in HBM before timing starts.

Default workload ("implicit-wing-2m"): the 3D transonic implicit case, plain
FP64 on the GPU -- a swept tapered wing (30 deg LE sweep, taper 0.56, AR 3.8,
10% closed-TE 4-digit section) in a half-span box with the far field at 10
chords, ~2M tetrahedra, Mach 0.84, alpha 3.06 deg, backward Euler with the
first-order Roe Jacobian in 5x5 blocks, ILU(0)-GMRES(30), adaptive CFL ramp
from 1 (capped at 1000).  The K timed steps follow W untimed steps of the ramp.
Besides them it measures the time to a 1e-8 drop of the density residual Linf
(sum_f F*A) from the uniform freestream for the plain FP64 form, the mixed
form (FP32 matrix, preconditioner and Krylov solve under the FP64 residual)
and the mixed form with the learned and with the aggregation AMG
preconditioner, and the lift/drag differences to the plain solve.

Per-kernel device times come from HIP events recorded on the solver's stream
during the timed region; the library also sums each launch's algorithmic
bytes (every array the kernel must touch, counted once), so roofline.achieved
= algorithmic bytes / device time of the dominant kernel.  The CPU baseline is
the oracle (a restatement of the reference loops, OpenMP) on a bounded sample.
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
KERNEL_NAMES = {"timestep": "k_timestep", "gradlim": "k_gradlim", "residual": "k_cell_residual",
                "faceflux": "k_face_flux",
                "jacobian": "k_face_jacobian", "assemble": "k_assemble", "spmv": "k_spmv",
                "precond": "k_ilu_lower+k_ilu_upper", "krylov": "k_mdot+k_orth_update",
                "factor": "k_ilu_factor_color", "update": "k_update", "learned": "learned preconditioner kernels"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# what the "precond" and "factor" kinds time in each implicit form
FORM_KERNEL_NAMES = {
    "mixed-learned": {"precond": "k_spmv<__half> (learned operator)", "factor": "learned operator build"},
    "mixed-amg": {"precond": "AMG V-cycle (amg:: kernels)", "factor": "AMG hierarchy build"},
}


def roofline(prof, names=None):
    names = {**KERNEL_NAMES, **(names or {})}
    kern = {k: v for k, v in prof.items() if v["ms"] > 0}
    dom = max(kern, key=lambda k: kern[k]["ms"])
    d = kern[dom]
    avg_ms = d["ms"] / d["launches"]
    name = names.get(dom, dom)
    tr = traffic_from_profiles(name)
    return {"bound": "hbm", "kernel": name, "achieved": round(d["GBps"], 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(d["GBps"] / HBM_PEAK_GBS, 4),
            "traffic": tr["bytes_per_launch"] if tr else None,
            "traffic_source": tr["source"] if tr else None,
            "alg_bytes_per_launch": round(d["bytes"] / d["launches"]),
            "avg_launch_us": round(avg_ms * 1e3, 2),
            "kernels": {names.get(k, k): {"total_ms": round(v["ms"], 3), "launches": v["launches"],
                                                  "avg_us": round(v["ms"] / v["launches"] * 1e3, 2),
                                                  "GBps": round(v["GBps"], 1)}
                        for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"])}}


# --------------------------------------------------------------------------
# implicit wing (default)
# --------------------------------------------------------------------------

# Pseudo-time schedule of the wing runs.  From the uniform freestream the
# first backward-Euler step at CFL 10 drives the thin cells under the leading
# edge to negative density (the oracle diverges the same way at 144k cells),
# so the ramp starts at CFL 1, as the reference's own 3D case does
# (COURANT_NUM = 1 in its ONERA M6 configuration), and grows with the
# reference's adaptive rule (SolverImplicit.cpp:530-550).  The cap is the
# largest at which the plain form (FP64 ILU(0)-GMRES(30) to rtol 5e-2) still
# converges on the 2M-cell systems: at CFL 5000 its solves stall and the
# residual freezes at a 1e-2 drop.
WING_CFL = (1.0, 1000.0, (0.5, 1.5))


TIMED_FORMS = {"plain": (64, "ilu0"), "mixed-ilu0": (32, "ilu0"), "mixed-learned": (32, "learned"),
               "mixed-amg": (32, "amg")}


def implicit_workload(args):
    from bes_amd import core, meshgen as mg
    from oracle import oracle as orc

    t0 = time.perf_counter()
    md = mg.swept_wing(*args.wing)
    spec = md.wing["spec"]
    alpha = 3.06
    mach = 0.84
    mvec = (mach * np.cos(np.radians(alpha)), 0.0, mach * np.sin(np.radians(alpha)))
    finf = orc.freestream_primitive(mvec, 3)
    U0 = np.tile(orc.conserved_from_primitive(finf, 3), (md.ncell, 1))
    log(f"mesh {md.ncell} tets in {time.perf_counter() - t0:.1f}s")
    cmin, cmax, ex = WING_CFL
    repeat = not args.no_repeat
    out = {}
    with core.Context(md) as c:
        c.set_freestream(finf)
        geo = c.geometry()
        log(f"context ready at {time.perf_counter() - t0:.1f}s")

        def setup(precision, precond):
            c.implicit_setup(precision, precond, restart=args.restart, maxiter=args.krylov_max, rtol=args.rtol)
            c.implicit_set_reorth(args.reorth)
            if precond == "learned":
                c.lp_init(20261202)
                c.lp_set_rebuild_interval(args.lp_rebuild)
            if precond == "amg":
                c.amg_set_params(relax="block_jacobi", rebuild_interval=args.amg_rebuild)

        def forces():
            U = c.get_state()
            return mg.force_coefficients(geo["S"], _centers(md, geo), geo["btype"], geo["owner"], U, mach,
                                         alpha, spec)[:2]

        def drop_run(name, precision, precond):
            setup(precision, precond)
            c.set_state(U0)
            c.synchronize()
            t = time.perf_counter()
            hs = [c.implicit_run(50, cmin, cmax, ex, stop_drop=args.drop, repeat_nv=repeat)]
            reached = len(hs[0]) < 50
            n = len(hs[0])
            r0 = hs[0][0, 5]
            while n < args.max_steps and not reached:
                h, reached = c.implicit_continue(min(50, args.max_steps - n))
                hs.append(h)
                n += len(h)
                log(f"  {name}: {n} steps {time.perf_counter() - t:.1f}s drop {h[-1, 5] / r0:.2e}")
            c.synchronize()
            wall = time.perf_counter() - t
            H = np.concatenate(hs)
            drop = float(H[-1, 5] / r0)
            cl, cd = forces()
            st = c.implicit_stats()
            log(f"  {name}: {len(H)} steps, drop {drop:.2e}, {wall:.2f}s, krylov {int(H[:, 11].sum())}, "
                f"CL {cl:.8f} CD {cd:.8f}")
            return dict(steps=len(H), wall_s=wall, drop=drop, reached=bool(drop <= args.drop),
                        krylov_total=int(H[:, 11].sum()), CL=cl, CD=cd, ilu0_fallbacks=st["lp_fallbacks"],
                        history_density_linf=[float(x) for x in H[::max(1, len(H) // 50), 5]])

        # K timed steps of the plain FP64 solve (or the form --timed-form names)
        # after W untimed steps of the ramp
        tprec, tprecond = TIMED_FORMS[args.timed_form]
        setup(tprec, tprecond)
        c.set_state(U0)
        c.implicit_run(args.warmup, cmin, cmax, ex, repeat_nv=repeat)
        c.synchronize()
        c.set_profiling(True)
        t = time.perf_counter()
        hist, _ = c.implicit_continue(args.steps)
        c.synchronize()
        wall = time.perf_counter() - t
        prof = c.profile()
        c.set_profiling(False)
        assert len(hist) == args.steps
        log(f"timed {args.steps} steps in {wall:.3f}s (CFL {hist[0, 10]:.1f} -> {hist[-1, 10]:.1f}, "
            f"{int(hist[:, 11].sum())} Krylov iterations)")
        value = md.ncell * args.steps / wall
        out["roofline"] = roofline(prof, FORM_KERNEL_NAMES.get(args.timed_form))
        out["timed_steps"] = {"cfl_first": float(hist[0, 10]), "cfl_last": float(hist[-1, 10]),
                              "krylov_iterations": int(hist[:, 11].sum())}
        if not args.no_drop:
            forms = args.forms.split(",")
            plain = drop_run("plain_fp64", 64, "ilu0")
            variants = {"plain_fp64_ilu0": plain}
            if "ilu0" in forms and not args.quick_drop:
                variants["mixed_fp32_ilu0"] = drop_run("mixed_fp32_ilu0", 32, "ilu0")
            if "learned" in forms:
                variants["mixed_fp32_learned"] = drop_run("mixed_fp32_learned", 32, "learned")
            if "amg" in forms:
                variants["mixed_fp32_amg"] = drop_run("mixed_fp32_amg", 32, "amg")
            for k, v in variants.items():
                if k != "plain_fp64_ilu0":
                    v["speedup_vs_plain"] = plain["wall_s"] / v["wall_s"]
                    v["dCL_vs_plain"] = abs(v["CL"] - plain["CL"])
                    v["dCD_vs_plain"] = abs(v["CD"] - plain["CD"])
            out["time_to_drop"] = {
                "target_drop": args.drop, "monitor": "Linf over cells of the density residual sum_f F*A",
                "schedule": {"cfl_min": cmin, "cfl_max": cmax, "exponents": list(ex),
                             "jacobian_block_repeat_nv": repeat},
                **variants,
                "learned_preconditioner": "network with its initial weights (blend 0.953 between block "
                                          "Jacobi and the first-order Neumann operator), rebuilt every %d "
                                          "solves; Adam training on the GPU lowers its loss but not yet its "
                                          "GMRES iterations" % args.lp_rebuild,
                "amg_preconditioner": "aggregation AMG (eps_strong 0.08, over_interp 1.5, coarsest <= 500 "
                                      "rows inverted) rebuilt every %d solves, V-cycle with damped (0.72) "
                                      "block Jacobi relaxation; sweeps from row-scaled FP16 copies"
                                      % args.amg_rebuild}
    cfgd = {"workload": "implicit-wing-2m" if args.wing == [96, 48, 72] else "implicit-wing",
            "mesh": f"swept tapered wing (LE sweep 30 deg, taper 0.56, AR 3.8, 10% closed-TE 4-digit "
                    f"section), half-span box, far field 10 chords, {args.wing[0]}x{args.wing[1]}x"
                    f"{args.wing[2]} lattice x 6 = {md.ncell} tetrahedra",
            "cells": md.ncell, "mach": mach, "alpha_deg": alpha,
            "scheme": "backward Euler, 2nd-order Roe residual (Green-Gauss, Venkatakrishnan K=5), "
                      "first-order Roe Jacobian in 5x5 blocks%s, %s-GMRES(%d) rtol %g, CFL %g->%g "
                      "exponents %s, local dt" % ("" if not repeat else " (nV-fold repeat as the reference)",
                                                  {"ilu0": "ILU(0)", "learned": "learned-operator",
                                                   "amg": "AMG"}[tprecond], args.restart, args.rtol, cmin, cmax,
                                                  ex),
            "precision": "FP64 throughout" if args.timed_form == "plain" else
                         "FP32 matrix, preconditioner and Krylov solve under the FP64 residual",
            "timed": f"{args.steps} steps of the {args.timed_form} form after {args.warmup} steps of the ramp"}
    res = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if tprec == 64 else "f64 residual, f32 solve",
           "data": "synthetic", "config": cfgd}
    res.update(out)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_implicit(args)
    return res


def _centers(md, geo):
    """Face centres of the reference face numbering (from the node coordinates)."""
    # faces: boundary faces keep file order at the front; their centres are what
    # the wing mask needs, internal faces get their owner centroid (unused).
    nf = geo["S"].shape[0]
    c = np.zeros((nf, md.ndim))
    nb = md.bfaces.shape[0]
    c[:nb] = md.nodes[md.bfaces].mean(axis=1)
    c[nb:] = geo["centroid"][geo["owner"][nb:]]
    return c


def cpu_baseline_implicit(args):
    """The oracle's implicit assembly (reference loops) on a sample, per cell."""
    from bes_amd import meshgen as mg
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    md = mg.swept_wing(24, 12, 18)
    mvec = (0.84 * np.cos(np.radians(3.06)), 0.0, 0.84 * np.sin(np.radians(3.06)))
    finf = orc.freestream_primitive(mvec, 3)
    U = orc.perturbed_state(md.ncell, 3, mvec, seed=1, amp=0.01)
    m = orc.Mesh(md)
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    nv = 5
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < args.cpu_seconds:
        rp, col, val, rhs, _ = m.implicit_assemble(U, 10.0, finf)
        # block CSR system, block-Jacobi preconditioned GMRES(30) as the step's solve
        A = sp.bsr_matrix((val, col, rp), shape=(md.ncell * nv, md.ncell * nv)).tocsr()
        Dinv = np.linalg.inv(val[rp[:-1]])
        M = spla.LinearOperator(A.shape, lambda r: np.einsum("nij,nj->ni", Dinv, r.reshape(-1, nv)).ravel())
        x, info = spla.gmres(A, rhs.ravel(), M=M, restart=30, maxiter=100, rtol=1e-3)
        U = U + x.reshape(U.shape)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": md.ncell * n / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{n} backward-Euler steps on a {md.ncell}-tet wing mesh, {dt:.1f} s: oracle "
                      f"assembly (reference face loop, Eigen-rule AD Jacobians, serial) + scipy "
                      f"block-Jacobi GMRES(30) rtol 1e-3 standing in for AMGCL (not buildable here: "
                      f"needs Boost); one thread (host threads available: {threads})"}


# --------------------------------------------------------------------------
# explicit workloads
# --------------------------------------------------------------------------

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
    cfg["cells"] = md.ncell
    out = {"metric": METRIC, "value": md.ncell * args.steps / wall, "unit": "cell-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic", "config": cfg, "roofline": roofline(prof),
           "residual_linf_last": [float(x) for x in linf[-1]]}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_explicit(args, nd)
    return out


RK3_ALPHA = (0.1481, 0.4, 1.0)   # Blazek's 3-stage coefficients for 2nd-order upwind schemes


def sweep_conditions(B, seed=20261202):
    """Seeded Mach uniform in [0.3, 0.85], angle of attack uniform in [-4, 8] deg."""
    rng = np.random.default_rng(seed)
    mach = rng.uniform(0.3, 0.85, B)
    aoa = rng.uniform(-4.0, 8.0, B)
    return mach, aoa


def sweep_workload(args):
    """B freestream conditions advancing together on one wing mesh (RK3)."""
    from bes_amd import core, meshgen as mg
    from oracle import oracle as orc

    t0 = time.perf_counter()
    md = mg.swept_wing(*args.sweep_wing)
    B = args.batch
    mach, aoa = sweep_conditions(B)
    finf = np.stack([orc.freestream_primitive((m * np.cos(np.radians(a)), 0.0, m * np.sin(np.radians(a))), 3)
                     for m, a in zip(mach, aoa)])
    U0 = np.stack([np.tile(orc.conserved_from_primitive(f, 3), (md.ncell, 1)) for f in finf])
    log(f"mesh {md.ncell} tets x {B} conditions in {time.perf_counter() - t0:.1f}s")
    cfl = 0.8
    with core.Context(md) as c:
        c.set_batch(B)
        c.set_freestream(finf)
        c.set_state(U0)
        del U0
        if args.graph:
            c.explicit_graph(True)
        c.explicit_run(args.warmup, cfl, alpha=RK3_ALPHA)
        c.synchronize()
        c.set_profiling(not args.graph)
        t = time.perf_counter()
        linf = c.explicit_run(args.steps, cfl, alpha=RK3_ALPHA)
        c.synchronize()
        wall = time.perf_counter() - t
        if args.graph:   # per-kernel figures from two further stream-launched steps
            c.explicit_graph(False)
            c.set_profiling(True)
            c.explicit_run(2, cfl, alpha=RK3_ALPHA)
        prof = c.profile()
        c.set_profiling(False)
    cfg = {"workload": f"explicit-sweep-{B}x{md.ncell // 1000}k",
           "mesh": f"swept tapered wing, {args.sweep_wing[0]}x{args.sweep_wing[1]}x{args.sweep_wing[2]} "
                   f"lattice x 6 = {md.ncell} tetrahedra", "cells": md.ncell, "conditions": B,
           "mach_range": [0.3, 0.85], "alpha_range_deg": [-4.0, 8.0],
           "scheme": f"RK3 alpha {RK3_ALPHA}, Roe, Green-Gauss, Venkatakrishnan K=5, local dt, CFL {cfl}",
           "launch": "one HIP graph per pseudo-time step" if args.graph else "stream launches"}
    out = {"metric": METRIC, "value": md.ncell * B * args.steps / wall, "unit": "cell-updates/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic", "config": cfg, "roofline": roofline(prof),
           "residual_linf_last_max": float(np.max(linf[-1]))}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_explicit(args, 3)
    return out


def implicit_sweep_workload(args):
    """The sweep's conditions as implicit mixed-precision solves (FP32 AMG-GMRES
    under the FP64 residual) to a residual drop, one condition after another on
    one shared mesh and solver context: the implicit kernels take one condition
    per launch, so the conditions share geometry, workspace and the AMG arena,
    not launches."""
    from bes_amd import core, meshgen as mg
    from oracle import oracle as orc

    t0 = time.perf_counter()
    md = mg.swept_wing(*args.sweep_wing)
    spec = md.wing["spec"]
    B = args.batch
    mach, aoa = sweep_conditions(B)
    cmin, cmax, ex = WING_CFL
    log(f"mesh {md.ncell} tets, {B} conditions, in {time.perf_counter() - t0:.1f}s")
    runs = []
    with core.Context(md) as c:
        geo = c.geometry()
        centers = _centers(md, geo)
        c.implicit_setup(32, "amg", restart=args.restart, maxiter=args.krylov_max, rtol=args.rtol)
        c.implicit_set_reorth(args.reorth)
        c.amg_set_params(relax="block_jacobi", rebuild_interval=args.amg_rebuild)
        t_all = time.perf_counter()
        total_steps = 0
        for q in range(B):
            mvec = (mach[q] * np.cos(np.radians(aoa[q])), 0.0, mach[q] * np.sin(np.radians(aoa[q])))
            finf = orc.freestream_primitive(mvec, 3)
            c.set_freestream(finf)
            c.set_state(np.tile(orc.conserved_from_primitive(finf, 3), (md.ncell, 1)))
            c.amg_set_params(relax="block_jacobi", rebuild_interval=args.amg_rebuild)   # fresh hierarchy
            c.synchronize()
            t = time.perf_counter()
            hs = [c.implicit_run(50, cmin, cmax, ex, stop_drop=args.drop, repeat_nv=False)]
            n, reached = len(hs[0]), len(hs[0]) < 50
            while n < args.max_steps and not reached:
                h, reached = c.implicit_continue(min(50, args.max_steps - n))
                hs.append(h)
                n += len(h)
            c.synchronize()
            wall = time.perf_counter() - t
            H = np.concatenate(hs)
            cl, cd = mg.force_coefficients(geo["S"], centers, geo["btype"], geo["owner"], c.get_state(),
                                           mach[q], aoa[q], spec)[:2]
            total_steps += len(H)
            runs.append(dict(mach=float(mach[q]), alpha_deg=float(aoa[q]), steps=len(H), wall_s=wall,
                             drop=float(H[-1, 5] / H[0, 5]), reached=bool(H[-1, 5] / H[0, 5] <= args.drop),
                             krylov_total=int(H[:, 11].sum()), CL=float(cl), CD=float(cd)))
            log(f"  condition {q}: M {mach[q]:.3f} alpha {aoa[q]:+.2f}: {len(H)} steps, {wall:.2f}s, "
                f"drop {runs[-1]['drop']:.1e}, CL {cl:.5f} CD {cd:.5f}")
        wall_all = time.perf_counter() - t_all
    cfg = {"workload": f"implicit-sweep-{B}x{md.ncell // 1000}k",
           "mesh": f"swept tapered wing, {args.sweep_wing[0]}x{args.sweep_wing[1]}x{args.sweep_wing[2]} "
                   f"lattice x 6 = {md.ncell} tetrahedra", "cells": md.ncell, "conditions": B,
           "mach_range": [0.3, 0.85], "alpha_range_deg": [-4.0, 8.0],
           "scheme": f"backward Euler, FP32 AMG-GMRES({args.restart}) rtol {args.rtol} under the FP64 residual, "
                     f"CFL {cmin}->{cmax}, to a {args.drop:g} drop per condition",
           "launch": "conditions one after another on one context (the implicit kernels take one "
                     "condition per launch)"}
    reached = sum(r["reached"] for r in runs)
    return {"metric": METRIC, "value": md.ncell * total_steps / wall_all, "unit": "cell-updates/s",
            "n_gpus": 1, "steps": total_steps, "warmup": 0, "ms_per_step": wall_all * 1e3 / total_steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 residual, f32 solve", "data": "synthetic", "config": cfg,
            "sweep": {"wall_s": wall_all, "conditions_reached": reached, "s_per_condition": wall_all / B,
                      "runs": runs}}


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


def copy_probe():
    """A plain device copy of known size: 1 GiB read and 1 GiB written per copy."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "copies": 5}))


def traffic_from_profiles(kernel):
    """HBM bytes per launch of a kernel as measured by the committed counter runs."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        table = json.load(fh)
    return table.get(kernel)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (implicit 30, explicit 20)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (implicit 100, explicit 3)")
    ap.add_argument("--workload", default="implicit-wing-2m",
                    choices=["implicit-wing-2m", "explicit-tetbox", "explicit-bump-40k", "explicit-sweep",
                             "implicit-sweep"])
    ap.add_argument("--batch", type=int, default=64, help="conditions of the explicit sweep")
    ap.add_argument("--sweep-wing", type=int, nargs=3, default=[60, 30, 46])
    ap.add_argument("--graph", action="store_true", help="replay each pseudo-time step as a HIP graph")
    ap.add_argument("--wing", type=int, nargs=3, default=[96, 48, 72])
    ap.add_argument("--box", type=int, nargs=3, default=[80, 64, 65])
    ap.add_argument("--max-steps", type=int, default=800)
    ap.add_argument("--krylov-max", type=int, default=300)
    ap.add_argument("--reorth", type=float, default=0.0, help="DGKS threshold (0: classical Gram-Schmidt once)")
    ap.add_argument("--no-repeat", action="store_true", default=True,
                    help="exact first-order Jacobian (the reference adds each block nV times)")
    ap.add_argument("--repeat", dest="no_repeat", action="store_false")
    ap.add_argument("--quick-drop", action="store_true", help="skip the mixed FP32 ILU(0) drop run")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device 5 times (calibrates HBM byte counters)")
    ap.add_argument("--timed-form", choices=sorted(TIMED_FORMS), default="plain",
                    help="form of the K timed implicit steps (the headline value is the plain form's)")
    ap.add_argument("--lp-rebuild", type=int, default=6,
                    help="learned operator rebuilt at every n-th solve of the mixed learned run")
    ap.add_argument("--amg-rebuild", type=int, default=6,
                    help="AMG hierarchy rebuilt at every n-th solve of the mixed AMG run")
    ap.add_argument("--restart", type=int, default=30, help="GMRES restart length of every implicit form")
    ap.add_argument("--drop", type=float, default=1e-8, help="residual drop of the time-to-drop runs")
    ap.add_argument("--forms", default="ilu0,learned,amg",
                    help="mixed forms of the time-to-drop runs besides the plain one (ilu0, learned, amg)")
    ap.add_argument("--rtol", type=float, default=5e-2,
                    help="GMRES relative tolerance of every form (the outer loop needs no tighter)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-drop", action="store_true")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    implicit = args.workload.startswith("implicit")
    if args.steps is None:
        args.steps = 30 if implicit else 20
    if args.warmup is None:
        args.warmup = 100 if implicit else 3
    if args.gpus != 1:
        print(json.dumps({"error": "this solver core runs on one GPU; the multi-GPU (partitioned mesh, "
                          "halo exchange) form is designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "implicit-sweep":
        out = implicit_sweep_workload(args)
    elif args.workload.startswith("implicit"):
        out = implicit_workload(args)
    elif args.workload == "explicit-sweep":
        out = sweep_workload(args)
    else:
        out = explicit_workload(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
