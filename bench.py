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


def next_cfl(cfl, lb_prev, lb_last, cmin, cmax, ex):
    """Courant number of the step after one run at cfl (SolverImplicit.cpp:530-550)."""
    e = ex[1] if lb_prev > lb_last else ex[0]
    return min(cmax, max(cmin, cfl * (lb_prev / lb_last) ** e))


def implicit_workload(args):
    from bes_amd import core, meshgen as mg

    t0 = time.perf_counter()
    md = mg.swept_wing(*args.wing)
    spec = md.wing["spec"]
    alpha = 3.06
    mach = 0.84
    finf = core.freestream_primitive(core.wing_freestream(mach, alpha), 3)
    U0 = np.tile(core.conserved_from_primitive(finf, 3), (md.ncell, 1))
    log(f"mesh {md.ncell} tets in {time.perf_counter() - t0:.1f}s")
    cmin, cmax, ex = WING_CFL
    repeat = not args.no_repeat
    out = {}
    with core.Context(md) as c:
        c.set_freestream(finf)
        geo = c.geometry()
        log(f"context ready at {time.perf_counter() - t0:.1f}s")

        tuned = {"coloring": None}

        def setup(precision, precond, params=None):
            kind = "learned" if precond.startswith("learned") else precond
            rtol = args.amg_rtol if kind == "amg" else args.lp_rtol if kind == "learned" else args.rtol
            c.implicit_setup(precision, kind, restart=args.restart, maxiter=args.krylov_max, rtol=rtol)
            c.implicit_set_reorth(args.reorth)
            if kind in ("ilu0", "learned") and tuned["coloring"]:   # the ILU(0) fallback of the learned form too
                c.implicit_set_coloring(tuned["coloring"])
            if kind == "learned":
                c.lp_init(20261202)
                if precond == "learned" and params is not None:
                    c.lp_set_params(params)
                elif precond in ("learned", "learned_untrained"):   # no trained weights: the neutral start
                    c.lp_neutral()
                c.lp_set_rebuild_interval(args.lp_rebuild)
            if precond == "amg":
                c.amg_set_params(relax="block_jacobi", rebuild_interval=args.amg_rebuild)
            c.set_limiter_freeze(args.limiter_freeze)   # 0: the reference's limiter at every pass

        def forces():
            U = c.get_state()
            return mg.force_coefficients(geo["S"], _centers(md, geo), geo["btype"], geo["owner"], U, mach,
                                         alpha, spec)[:2]

        frozen = {"phi": None}   # the plain run's frozen limiter, shared by the other forms

        def drop_run(name, precision, precond, params=None):
            setup(precision, precond, params)
            c.set_state(U0)
            if frozen["phi"] is not None:
                c.freeze_limiter_values(frozen["phi"])
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
                        extra_restarts=st.get("extra_restarts", 0),
                        history_density_linf=[float(x) for x in H[::max(1, len(H) // 50), 5]])

        # The timed steps run at the Courant-number cap: the ramp from CFL 1
        # runs untimed until a step has used the cap (its first steps take a
        # few Krylov iterations each and do not represent the solve), then W
        # more untimed steps, then exactly K timed steps.
        tprec, tprecond = TIMED_FORMS[args.timed_form]
        setup(tprec, tprecond)
        c.set_state(U0)
        ramp = [c.implicit_run(2, cmin, cmax, ex, repeat_nv=repeat)]
        nramp = 2
        while ramp[-1][-1, 10] < cmax and nramp < args.max_ramp:
            h, _ = c.implicit_continue(5)
            ramp.append(h)
            nramp += len(h)
        if args.warmup > 0:
            h, _ = c.implicit_continue(args.warmup)
            ramp.append(h)
        R = np.concatenate(ramp)
        log(f"ramp: {nramp} steps to CFL {R[nramp - 1, 10]:.1f}, then {args.warmup} warmup steps")
        if args.autotune and tprecond == "ilu0":   # the tuner covers the ILU(0) form's kernels
            # block sizes and the ILU(0) colouring, timed on this mesh's system at the CFL cap
            ta = time.perf_counter()
            c.implicit_assemble(R[-1, 10], repeat_nv=repeat)
            bs, col = c.autotune(implicit=True, coloring=tprecond == "ilu0")
            tuned["coloring"] = col
            out["autotune"] = {"block_size_us": bs, "ilu0_coloring": col,
                               "ncolors": c.implicit_ncolors() if tprecond == "ilu0" else None,
                               "wall_s": time.perf_counter() - ta}
            log(f"autotune {time.perf_counter() - ta:.1f}s: {bs}, coloring {col}")
        # state, Courant number and monitors at the start of the timed steps
        # (the CPU baseline continues from exactly this point)
        U_start = c.get_state()
        loop = {"cfl": next_cfl(R[-1, 10], R[-2, 0], R[-1, 0], cmin, cmax, ex), "lb_old": R[-1, :5].copy()}
        c.synchronize()
        c.set_profiling(True)
        # each implicit step ends with a host wait (the solve's residual
        # checks and the monitors), so per-step host times are device times
        step_s = []
        hs = []
        t = time.perf_counter()
        for _ in range(args.steps):
            ts = time.perf_counter()
            hs.append(c.implicit_continue(1)[0])
            step_s.append(time.perf_counter() - ts)
        c.synchronize()
        wall = time.perf_counter() - t
        prof = c.profile()
        c.set_profiling(False)
        hist = np.concatenate(hs)
        assert len(hist) == args.steps
        log(f"timed {args.steps} steps in {wall:.3f}s (CFL {hist[0, 10]:.1f} -> {hist[-1, 10]:.1f}, "
            f"{int(hist[:, 11].sum())} Krylov iterations)")
        value = md.ncell * args.steps / wall
        out["roofline"] = roofline(prof, FORM_KERNEL_NAMES.get(args.timed_form))
        out["timed_steps"] = {"ramp_steps_to_cap": nramp, "warmup_steps": args.warmup,
                              "cfl_first": float(hist[0, 10]), "cfl_last": float(hist[-1, 10]),
                              "cfl_min": float(hist[:, 10].min()),
                              "krylov_iterations": int(hist[:, 11].sum()),
                              "krylov_per_step": [int(x) for x in hist[:, 11]],
                              "ms_per_step_each": [round(x * 1e3, 2) for x in step_s],
                              "density_drop_at_start": float(R[-1, 5] / R[0, 5])}
        if not args.no_drop:
            forms = args.forms.split(",")
            plain = drop_run("plain_fp64", 64, "ilu0")
            if args.limiter_freeze > 0:
                # every other form converges the discretisation the plain run
                # froze (its limiter from the step its drop reached the level)
                plain["limiter_frozen"] = bool(c.limiter_frozen()[0])
                frozen["phi"] = c.gradient_limiter()[1]
            variants = {"plain_fp64_ilu0": plain}
            if "ilu0" in forms and not args.quick_drop:
                variants["mixed_fp32_ilu0"] = drop_run("mixed_fp32_ilu0", 32, "ilu0")
            if "learned" in forms:
                params, training = train_learned(args, md)
                out["learned_training"] = training
                variants["mixed_fp32_learned"] = drop_run("mixed_fp32_learned", 32, "learned", params)
            if "learned_untrained" in forms:
                variants["mixed_fp32_learned_untrained"] = drop_run("mixed_fp32_learned_untrained", 32,
                                                                    "learned_untrained")
            if "learned_handset" in forms:
                variants["mixed_fp32_learned_handset"] = drop_run("mixed_fp32_learned_handset", 32,
                                                                  "learned_handset")
            if "amg" in forms:
                variants["mixed_fp32_amg"] = drop_run("mixed_fp32_amg", 32, "amg")
            plain["cell_updates_per_s_whole_run"] = md.ncell * plain["steps"] / plain["wall_s"]
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
                "learned_preconditioner": LEARNED_NOTE % args.lp_rebuild,
                "amg_preconditioner": "aggregation AMG (eps_strong 0.08, over_interp 1.5, coarsest <= 500 "
                                      "rows inverted) rebuilt every %d solves, V-cycle with damped (0.72) "
                                      "block Jacobi relaxation; sweeps from row-scaled FP16 copies; GMRES "
                                      "rtol %g" % (args.amg_rebuild, args.amg_rtol)}
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
            "timed": f"{args.steps} steps of the {args.timed_form} form at the CFL cap: the ramp ran "
                     f"{nramp} untimed steps until a step used CFL {cmax:g}, then {args.warmup} untimed steps"}
    if args.limiter_freeze > 0:
        cfgd["limiter_freeze"] = (f"the plain run fixes the limiter once its density residual drop reaches "
                                  f"{args.limiter_freeze:g}; the other forms use that frozen limiter from their "
                                  f"first step, so all converge one discretisation (no reference counterpart)")
    res = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if tprec == 64 else "f64 residual, f32 solve",
           "data": "synthetic", "config": cfgd}
    res.update(out)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_implicit(args, md, finf, U_start, loop, hist, step_s)
        if not args.no_config1:
            res["config1_explicit_bump_40k"] = config1_explicit(args)
    return res


LEARNED_NOTE = ("network trained on the GPU from a neutral start (every blend 1/2) by online Adam steps on the "
                "systems of an implicit run on the 0.5M-tet wing (learned_training), then used unchanged on the "
                "2M wing, rebuilt every %d solves; _untrained: the neutral network; _handset: the hand-set "
                "decoder bias (blend 0.953)")


def train_learned(args, md_eval):
    """Train the learned preconditioner on the GPU, from a neutral start, on
    the Jacobians and residuals an implicit run generates on the 0.5M-tet wing
    (a mesh the timed 2M runs never see): a few Adam steps on each assembled
    system (bes_lp_set_online)."""
    from bes_amd import core, meshgen as mg
    md = mg.swept_wing(*args.sweep_wing)
    finf = core.freestream_primitive(core.wing_freestream(0.84, 3.06), 3)
    cmin, cmax, ex = WING_CFL
    t0 = time.perf_counter()
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(np.tile(core.conserved_from_primitive(finf, 3), (md.ncell, 1)))
        c.implicit_setup(32, "learned", restart=args.restart, maxiter=args.krylov_max, rtol=args.rtol)
        c.implicit_set_reorth(args.reorth)
        c.lp_init(20261202)
        c.lp_neutral()
        c.lp_set_online(args.lp_train_steps, lr=args.lp_lr, nprobe=4, nrich=3)
        h = c.implicit_run(args.lp_train_pseudo, cmin, cmax, ex, stop_drop=1e-6, repeat_nv=False)
        params = c.lp_params()
        loss = c.lp_online_loss()
    wall = time.perf_counter() - t0
    log(f"learned training: {len(h)} pseudo-time steps x {args.lp_train_steps} Adam steps on {md.ncell} tets, "
        f"{wall:.1f}s, last loss {loss:.3f}")
    return params, {"mesh_cells": md.ncell, "pseudo_time_steps": len(h), "adam_steps_per_system":
                    args.lp_train_steps, "systems": len(h), "learning_rate": args.lp_lr, "last_loss": loss,
                    "wall_s": wall, "krylov_per_step_first_last": [int(h[0, 11]), int(h[-1, 11])]}


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


try:   # the process's CPU set before an OpenMP runtime binds the main thread to one place
    AFFINITY_AT_START = len(os.sched_getaffinity(0))
except AttributeError:
    AFFINITY_AT_START = None


def host_info(threads):
    """CPU model, core counts and OpenMP settings of the host the baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": AFFINITY_AT_START, "omp_threads": threads,
            "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"), "OMP_PLACES": os.environ.get("OMP_PLACES"),
            "compiler": "g++ " + ORACLE_FLAGS}


ORACLE_FLAGS = "-std=c++17 -O3 -fPIC -fopenmp -ffp-contract=off (oracle/Makefile)"


def omp_threads():
    """Threads of the CPU baseline: the host's CPU share (OMP_NUM_THREADS, else the affinity mask)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return AFFINITY_AT_START or os.cpu_count() or 1


def cpu_baseline_implicit(args, md, finf, U_start, loop, gpu_hist, gpu_step_s):
    """The oracle's implicit loop (OpenMP assembly with atomics, multicolour block
    ILU(0)-GMRES(30) at the same tolerance) continuing from the GPU's state at
    the start of its timed steps: same mesh, state, Courant number and solver
    settings.  Steps are timed one at a time until cpu_seconds have passed."""
    from oracle import oracle as orc
    threads = omp_threads()
    t0 = time.perf_counter()
    m = orc.Mesh(md)
    s = orc.Implicit(m, restart=args.restart, maxiter=args.krylov_max, rtol=args.rtol, nthreads=threads)
    cmin, cmax, ex = WING_CFL
    s.start(cmin, cmax, ex, repeat_nv=not args.no_repeat)
    s.resume(loop["cfl"], loop["lb_old"], -1.0)
    U = np.ascontiguousarray(U_start, dtype=np.float64)
    log(f"cpu baseline: oracle set-up {time.perf_counter() - t0:.1f}s, {threads} threads")
    hs = []
    t = time.perf_counter()
    while True:
        hs.append(s.steps(U, 1, finf))
        if time.perf_counter() - t > args.cpu_seconds or len(hs) >= args.steps:
            break
    wall = time.perf_counter() - t
    H = np.concatenate(hs)
    n = len(H)
    log(f"cpu baseline: {n} steps in {wall:.1f}s, Krylov {H[:, 11].tolist()}")
    return {"value": md.ncell * n / wall, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} backward-Euler step(s) of the same {md.ncell}-tet wing from the GPU's state at the "
                      f"start of its timed steps (CFL {H[0, 10]:g}), {wall:.1f} s: the oracle's implicit loop "
                      f"(face-parallel assembly with atomic adds as SolverImplicit.cpp:368, Eigen-rule AD "
                      f"Jacobians, multicolour block ILU(0)-GMRES({args.restart}) rtol {args.rtol:g} in FP64; the "
                      f"reference's AMGCL solve cannot be built here) on {threads} OpenMP threads",
            "ms_per_step": wall * 1e3 / n, "krylov_per_step": [int(x) for x in H[:, 11]],
            "gpu_same_steps": {"ms_per_step": sum(gpu_step_s[:n]) * 1e3 / n,
                               "krylov_per_step": [int(x) for x in gpu_hist[:n, 11]]},
            "gpu_over_cpu_same_steps": wall / sum(gpu_step_s[:n]),
            "host": host_info(threads)}


def config1_explicit(args):
    """Config 1 in full (2D bump channel, 40k triangles, RK4 at CFL 0.8, 2000
    iterations from the uniform freestream) on the GPU and with the oracle's
    OpenMP loop on the host cores: wall time of each, and the final residuals."""
    from bes_amd import core, meshgen as mg
    from oracle import oracle as orc
    md = mg.bump_channel(200, 100)
    finf = core.freestream_primitive((0.5, 0.0), 2)
    U0 = np.tile(core.conserved_from_primitive(finf, 2), (md.ncell, 1))
    niter = args.config1_iters
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U0)
        c.explicit_graph(True)
        c.explicit_run(2, 0.8)   # graph capture outside the timing
        c.set_state(U0)
        c.synchronize()
        t = time.perf_counter()
        lg = c.explicit_run(niter, 0.8)
        c.synchronize()
        wall_gpu = time.perf_counter() - t
        Ug = c.get_state()
        # the persistent form: all steps in one cooperative launch
        c.explicit_graph(False)
        c.explicit_persistent(True)
        c.explicit_run(2, 0.8)
        c.set_state(U0)
        c.synchronize()
        t = time.perf_counter()
        lp = c.explicit_run(niter, 0.8)
        c.synchronize()
        wall_pers = time.perf_counter() - t
        same = bool(np.array_equal(c.get_state(), Ug) and np.array_equal(lp, lg))
        c.explicit_persistent(False)
    threads = omp_threads()
    m = orc.Mesh(md)
    t = time.perf_counter()
    Uc, lc = m.explicit(U0, niter, 0.8, finf, nthreads=threads)
    wall_cpu = time.perf_counter() - t
    log(f"config 1: GPU {wall_gpu:.2f}s (graph), {wall_pers:.2f}s (persistent), CPU {wall_cpu:.2f}s "
        f"for {niter} iterations")
    return {"workload": "explicit-bump-40k", "cells": md.ncell, "iterations": niter,
            "scheme": "RK4 (4-stage), Roe, Green-Gauss, Venkatakrishnan K=5, local dt, CFL 0.8, FP64",
            "gpu": {"wall_s": wall_gpu, "cell_updates_per_s": md.ncell * niter / wall_gpu,
                    "launch": "one HIP graph per pseudo-time step",
                    "density_linf_drop": float(lg[-1, 0, 0] / lg[0, 0, 0]) if lg.ndim == 3 else
                    float(lg[-1, 0] / lg[0, 0])},
            "gpu_persistent": {"wall_s": wall_pers, "cell_updates_per_s": md.ncell * niter / wall_pers,
                               "launch": "one cooperative launch for all iterations",
                               "bitwise_equal_to_graph": same},
            "cpu": {"wall_s": wall_cpu, "cell_updates_per_s": md.ncell * niter / wall_cpu, "cores": threads,
                    "kind": "port", "density_linf_drop": float(lc[-1, 0] / lc[0, 0])},
            "gpu_over_cpu": wall_cpu / wall_gpu,
            "max_abs_state_difference": float(np.abs(Ug - Uc).max())}


# --------------------------------------------------------------------------
# explicit workloads
# --------------------------------------------------------------------------

def explicit_workload(args):
    from bes_amd import core, meshgen as mg

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
    finf = core.freestream_primitive(mach, nd)
    U0 = np.tile(core.conserved_from_primitive(finf, nd), (md.ncell, 1))
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

    t0 = time.perf_counter()
    md = mg.swept_wing(*args.sweep_wing)
    B = args.batch
    mach, aoa = sweep_conditions(B)
    finf = np.stack([core.freestream_primitive(core.wing_freestream(m, a), 3) for m, a in zip(mach, aoa)])
    U0 = np.stack([np.tile(core.conserved_from_primitive(f, 3), (md.ncell, 1)) for f in finf])
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
    """The sweep's conditions as implicit mixed-precision solves (FP32 ILU(0)-GMRES
    under the FP64 residual) to a residual drop, args.batch conditions per
    batched context (one launch per kernel for the whole batch), each condition
    with its own Courant number, Krylov stopping point and stop."""
    from bes_amd import core, meshgen as mg

    t0 = time.perf_counter()
    md = mg.swept_wing(*args.sweep_wing)
    spec = md.wing["spec"]
    B = args.isweep_batch
    mach, aoa = sweep_conditions(args.sweep_conditions)
    # a batch iterates until its hardest condition converges, so conditions of
    # similar stiffness share a batch: sorted by Mach number
    order = np.argsort(mach, kind="stable")
    mach, aoa = mach[order], aoa[order]
    # the FP32 ILU(0) solves of some conditions stagnate at the wing's CFL cap of
    # 1000 (GMRES reaches maxiter every step and the residual stops); the sweep
    # caps the Courant number lower
    cmin, _, ex = WING_CFL
    cmax = args.sweep_cfl_max
    log(f"mesh {md.ncell} tets, {len(mach)} conditions in batches of {B}, in {time.perf_counter() - t0:.1f}s")
    runs = []
    wall_all = 0.0
    total_cu = 0
    with core.Context(md) as c:
        geo = c.geometry()
        centers = _centers(md, geo)
        c.set_batch(B)
        for q0 in range(0, len(mach), B):
            qs = list(range(q0, q0 + B))
            P = np.stack([core.freestream_primitive(core.wing_freestream(mach[q], aoa[q]), 3) for q in qs])
            c.set_freestream(P)
            c.set_state(np.stack([np.tile(core.conserved_from_primitive(p, 3), (md.ncell, 1)) for p in P]))
            c.implicit_setup(32, "ilu0", restart=args.restart, maxiter=args.krylov_max, rtol=args.rtol)
            c.implicit_set_reorth(args.reorth)
            c.implicit_set_coloring("mesh")   # the order the autotuner picks on the 2M wing
            # each condition fixes its limiter once its own drop reaches the freeze
            # level: at some Mach numbers and angles the limiter of a few cells
            # cycles and the condition's residual levels off otherwise
            c.set_limiter_freeze(args.limiter_freeze if args.limiter_freeze > 0 else 1e-2)
            c.synchronize()
            t = time.perf_counter()
            c.implicit_run(10, cmin, cmax, ex, stop_drop=args.drop, repeat_nv=False)
            st = c.implicit_batch_state()
            while st["steps"].max() < args.max_steps and not st["reached"].all():
                c.implicit_continue(min(10, args.max_steps - int(st["steps"].max())))
                st = c.implicit_batch_state()
                log(f"  batch {q0 // B}: {int(st['steps'].max())} steps, {int(st['reached'].sum())}/{B} reached, "
                    f"{time.perf_counter() - t:.1f}s")
            c.synchronize()
            wall = time.perf_counter() - t
            wall_all += wall
            U = c.get_state()
            for k, q in enumerate(qs):
                cl, cd = mg.force_coefficients(geo["S"], centers, geo["btype"], geo["owner"], U[k], mach[q], aoa[q],
                                               spec)[:2]
                runs.append(dict(mach=float(mach[q]), alpha_deg=float(aoa[q]), steps=int(st["steps"][k]),
                                 reached=bool(st["reached"][k]), CL=float(cl), CD=float(cd)))
            total_cu += md.ncell * int(st["steps"].sum())
            log(f"  batch {q0 // B}: {wall:.1f}s, steps {st['steps'].tolist()}")
    cfg = {"workload": f"implicit-sweep-{len(mach)}x{md.ncell // 1000}k",
           "mesh": f"swept tapered wing, {args.sweep_wing[0]}x{args.sweep_wing[1]}x{args.sweep_wing[2]} "
                   f"lattice x 6 = {md.ncell} tetrahedra", "cells": md.ncell, "conditions": len(mach),
           "mach_range": [0.3, 0.85], "alpha_range_deg": [-4.0, 8.0],
           "scheme": f"backward Euler, FP32 ILU(0)-GMRES({args.restart}) rtol {args.rtol} under the FP64 residual, "
                     f"CFL {cmin}->{cmax}, to a {args.drop:g} drop per condition, each condition's limiter frozen "
                     f"when its drop reaches {args.limiter_freeze if args.limiter_freeze > 0 else 1e-2:g}",
           "launch": f"batches of {B} conditions sorted by Mach number, one launch per kernel for a whole batch; "
                     f"ILU(0) in the mesh-order colouring"}
    reached = sum(r["reached"] for r in runs)
    amg = None
    if args.isweep_amg:
        amg = implicit_sweep_amg(args, md, mach, aoa, freeze=args.limiter_freeze if args.limiter_freeze > 0 else 1e-2)
    return {"metric": METRIC, "value": total_cu / wall_all, "unit": "cell-updates/s",
            "n_gpus": 1, "steps": int(sum(r["steps"] for r in runs)), "warmup": 0,
            "ms_per_step": wall_all * 1e3 / max(1, sum(r["steps"] for r in runs)) * B,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 residual, f32 solve", "data": "synthetic", "config": cfg,
            "sweep": {"wall_s": wall_all, "conditions_reached": reached, "s_per_condition": wall_all / len(mach),
                      "runs": runs},
            "sweep_amg_one_at_a_time": amg}


def implicit_sweep_amg(args, md, mach, aoa, freeze):
    """The same conditions one at a time on one context with the fastest
    single-condition form (FP32 aggregation AMG-GMRES, the wing's CFL cap):
    the comparison for the batched ILU(0) sweep, whose batches iterate until
    their hardest condition converges."""
    from bes_amd import core, meshgen as mg
    cmin, cmax, ex = WING_CFL
    runs, wall_all = [], 0.0
    with core.Context(md) as c:
        geo = c.geometry()
        centers = _centers(md, geo)
        c.implicit_setup(32, "amg", restart=args.restart, maxiter=args.krylov_max, rtol=args.amg_rtol)
        c.implicit_set_reorth(args.reorth)
        for q in range(len(mach)):
            # new parameters force a fresh hierarchy: no lagged one from the last condition
            c.amg_set_params(relax="block_jacobi", rebuild_interval=args.amg_rebuild)
            P = core.freestream_primitive(core.wing_freestream(mach[q], aoa[q]), 3)
            c.set_freestream(P)
            c.set_state(np.tile(core.conserved_from_primitive(P, 3), (md.ncell, 1)))
            c.set_limiter_freeze(freeze)
            c.synchronize()
            t = time.perf_counter()
            h = c.implicit_run(args.max_steps, cmin, cmax, ex, stop_drop=args.drop, repeat_nv=False)
            c.synchronize()
            wall = time.perf_counter() - t
            wall_all += wall
            cl, cd = mg.force_coefficients(geo["S"], centers, geo["btype"], geo["owner"], c.get_state(), mach[q],
                                           aoa[q], md.wing["spec"])[:2]
            reached = bool(h[-1, 5] <= args.drop * h[0, 5])
            runs.append(dict(mach=float(mach[q]), alpha_deg=float(aoa[q]), steps=len(h), reached=reached,
                             wall_s=wall, CL=float(cl), CD=float(cd)))
            log(f"  amg condition {q}: M {mach[q]:.3f} alpha {aoa[q]:.2f}, {len(h)} steps, {wall:.2f}s, "
                f"reached {reached}")
    return {"wall_s": wall_all, "conditions_reached": sum(r["reached"] for r in runs),
            "s_per_condition": wall_all / len(mach),
            "cell_updates_per_s": md.ncell * sum(r["steps"] for r in runs) / wall_all,
            "scheme": f"FP32 AMG-GMRES({args.restart}) rtol {args.amg_rtol}, hierarchy rebuilt every "
                      f"{args.amg_rebuild} solves, CFL {cmin}->{cmax}, limiter frozen at a {freeze:g} drop",
            "runs": runs}


def implicit_bump_workload(args):
    """Mixed precision on the 2D bump channel (40,000 triangles, M 0.5), the
    implicit loop to a 1e-10 drop: the plain FP64 form and the mixed forms
    (FP32 Jacobian, ILU(0) or block-Jacobi preconditioner and GMRES under the
    FP64 residual) on the GPU, and the oracle's FP64 loop on the host cores;
    outer step counts and the bump's force coefficients against the oracle's."""
    from bes_amd import core, meshgen as mg
    from oracle import oracle as orc

    md = mg.bump_channel(200, 100)
    nx = 200   # the lower wall's edges are the first boundary faces
    mach = (0.5, 0.0)
    finf = core.freestream_primitive(mach, 2)
    U0 = np.tile(core.conserved_from_primitive(finf, 2), (md.ncell, 1))
    cmin, cmax, ex = WING_CFL
    drop = args.drop if args.drop < 1e-8 else 1e-10
    gamma = 1.4
    qL = 0.5 * (mach[0] ** 2 + mach[1] ** 2) * 1.0   # rho = 1, bump chord 1

    def forces(geo, U):
        # pressure force on the lower wall, gauge pressure p - p_inf (the flat
        # parts of the wall then add nothing), over dynamic pressure x chord
        o = geo["owner"][:nx]
        rho = U[o, 0]
        p = (gamma - 1) * (U[o, 3] - 0.5 * (U[o, 1] ** 2 + U[o, 2] ** 2) / rho) - finf[3]
        f = (p[:, None] * geo["S"][:nx]).sum(0) / qL
        return float(f[0]), float(f[1])

    # the limiter of this mesh cycles in a few cells near the bump's ends and the
    # residual levels off at a 2e-2 drop (the oracle's too): the plain run fixes
    # it once its drop reaches the freeze level, and every other run (the mixed
    # forms and the oracle) uses that frozen limiter from its first step, so all
    # converge one discretisation
    freeze = args.limiter_freeze if args.limiter_freeze > 0 else 1e-2
    phi = None
    out = {}
    with core.Context(md) as c:
        geo = c.geometry()
        c.set_freestream(finf)
        for name, prec, precond in (("plain_fp64_ilu0", 64, "ilu0"), ("mixed_fp32_ilu0", 32, "ilu0"),
                                    ("mixed_fp32_block_jacobi", 32, "jacobi")):
            c.set_state(U0)
            c.implicit_setup(prec, precond, restart=args.restart, maxiter=args.krylov_max, rtol=args.rtol)
            c.implicit_set_reorth(args.reorth)
            if phi is None:
                c.set_limiter_freeze(freeze)
            else:
                c.freeze_limiter_values(phi)
            c.synchronize()
            t = time.perf_counter()
            H = c.implicit_run(args.max_steps, cmin, cmax, ex, stop_drop=drop, repeat_nv=False)
            c.synchronize()
            wall = time.perf_counter() - t
            cx, cy = forces(geo, c.get_state())
            out[name] = dict(steps=len(H), wall_s=wall, drop=float(H[-1, 4] / H[0, 4]),
                             krylov_total=int(H[:, 9].sum()), Cx=cx, Cy=cy)
            if phi is None:
                out[name]["limiter_frozen"] = bool(c.limiter_frozen()[0])
                phi = c.gradient_limiter()[1]
            log(f"  bump {name}: {len(H)} steps, {wall:.2f}s, drop {out[name]['drop']:.2e}, Cx {cx:.10f} Cy {cy:.10f}")
    threads = omp_threads()
    om = orc.Mesh(md)
    s = orc.Implicit(om, restart=args.restart, maxiter=args.krylov_max, rtol=args.rtol, nthreads=threads)
    s.freeze_limiter(phi)
    s.start(cmin, cmax, ex, stop_drop=drop)
    U = np.ascontiguousarray(U0.copy())
    t = time.perf_counter()
    H = s.steps(U, args.max_steps, finf)
    wall = time.perf_counter() - t
    fo = om.faces()
    geo_o = {"owner": fo["owner"], "S": fo["S"]}
    cx, cy = forces(geo_o, U)
    oracle = dict(steps=len(H), wall_s=wall, drop=float(H[-1, 4] / H[0, 4]), krylov_total=int(H[:, 9].sum()),
                  Cx=cx, Cy=cy, cores=threads)
    log(f"  bump oracle FP64: {len(H)} steps, {wall:.2f}s, drop {oracle['drop']:.2e}, Cx {cx:.10f} Cy {cy:.10f}")
    for k, v in out.items():
        v["dCx_vs_oracle"] = abs(v["Cx"] - cx)
        v["dCy_vs_oracle"] = abs(v["Cy"] - cy)
        if k != "plain_fp64_ilu0":
            v["speedup_vs_plain"] = out["plain_fp64_ilu0"]["wall_s"] / v["wall_s"]
    m = out["mixed_fp32_ilu0"]
    cfg = {"workload": "implicit-bump-40k-mixed", "mesh": "2D channel, 10% circular-arc bump, 200x100 quads "
           "split to 40,000 triangles", "cells": md.ncell, "mach": 0.5,
           "scheme": f"backward Euler, 2nd-order Roe residual, first-order Roe Jacobian in 4x4 blocks, "
                     f"GMRES({args.restart}) rtol {args.rtol}, CFL {cmin}->{cmax} exponents {ex}, local dt, "
                     f"to a {drop:g} drop",
           "limiter_freeze": f"the plain run fixes the limiter when its drop reaches {freeze:g}; the mixed "
                             f"forms and the oracle use that frozen limiter from their first step",
           "precision": "the value is the mixed FP32 ILU(0) form's: FP32 Jacobian, ILU(0) and Krylov vectors "
                        "under the FP64 residual"}
    return {"metric": METRIC, "value": md.ncell * m["steps"] / m["wall_s"], "unit": "cell-updates/s",
            "n_gpus": 1, "steps": m["steps"], "warmup": 0, "ms_per_step": m["wall_s"] * 1e3 / m["steps"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 residual, f32 solve",
            "data": "synthetic", "config": cfg,
            "time_to_drop": {**out, "oracle_fp64": oracle,
                             "force": "pressure force on the lower wall (the bump), gauge pressure of the wall "
                                      "cells, over dynamic pressure x bump chord: Cx drag, Cy lift"},
            "cpu_baseline": {"value": md.ncell * oracle["steps"] / oracle["wall_s"], "unit": "cell-updates/s",
                             "cores": threads, "kind": "port",
                             "sample": f"the oracle's FP64 implicit loop to the same drop ({oracle['steps']} steps)"}}


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
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    y.copy_(x)
    e0.record()
    for _ in range(5):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    s = x.sum()
    e0.record()
    for _ in range(5):
        s = x.sum()
    e1.record()
    torch.cuda.synchronize()
    ms_read = e0.elapsed_time(e1) / 5
    # each copy reads and writes n*4 bytes, each sum reads them: the stream
    # rates a kernel can reach in practice
    print(json.dumps({"copy_bytes": n * 4, "copies": 6, "ms_per_copy": ms,
                      "GBps_read_plus_write": 2 * n * 4 / (ms * 1e-3) / 1e9,
                      "ms_per_sum": ms_read, "GBps_read": n * 4 / (ms_read * 1e-3) / 1e9}))


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
    ap.add_argument("--steps", type=int, default=None, help="timed steps (implicit 20, explicit 20)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps before the timed ones (implicit: after the ramp has reached the CFL "
                         "cap; default 5), explicit 3")
    ap.add_argument("--workload", default="implicit-wing-2m",
                    choices=["implicit-wing-2m", "explicit-tetbox", "explicit-bump-40k", "explicit-sweep",
                             "implicit-sweep", "implicit-bump"])
    ap.add_argument("--batch", type=int, default=64, help="conditions per batch of the sweeps")
    ap.add_argument("--sweep-conditions", type=int, default=64, help="conditions of the implicit sweep")
    ap.add_argument("--isweep-batch", type=int, default=16, help="conditions per batch of the implicit sweep")
    ap.add_argument("--sweep-cfl-max", type=float, default=200.0, help="CFL cap of the implicit sweep")
    ap.add_argument("--isweep-amg", action="store_true",
                    help="implicit sweep: also run the conditions one at a time with the AMG form")
    ap.add_argument("--sweep-wing", type=int, nargs=3, default=[60, 30, 46])
    ap.add_argument("--graph", action="store_true", help="replay each pseudo-time step as a HIP graph")
    ap.add_argument("--wing", type=int, nargs=3, default=[96, 48, 72])
    ap.add_argument("--box", type=int, nargs=3, default=[80, 64, 65])
    ap.add_argument("--max-steps", type=int, default=800)
    ap.add_argument("--limiter-freeze", type=float, default=0.0,
                    help="fix the limiter once the density residual drop reaches this level (0: never)")
    ap.add_argument("--krylov-max", type=int, default=300)
    ap.add_argument("--reorth", type=float, default=0.0, help="DGKS threshold (0: classical Gram-Schmidt once)")
    ap.add_argument("--no-repeat", action="store_true", default=True,
                    help="exact first-order Jacobian (the reference adds each block nV times)")
    ap.add_argument("--repeat", dest="no_repeat", action="store_false")
    ap.add_argument("--quick-drop", action="store_true", help="skip the mixed FP32 ILU(0) drop run")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device 6 times (calibrates HBM byte counters; prints the copy rate)")
    ap.add_argument("--timed-form", choices=sorted(TIMED_FORMS), default="plain",
                    help="form of the K timed implicit steps (the headline value is the plain form's)")
    ap.add_argument("--lp-rebuild", type=int, default=6,
                    help="learned operator rebuilt at every n-th solve of the mixed learned run")
    ap.add_argument("--lp-rtol", type=float, default=5e-2,
                    help="GMRES relative tolerance of the learned forms")
    ap.add_argument("--amg-rebuild", type=int, default=10,
                    help="AMG hierarchy rebuilt at every n-th solve of the mixed AMG run")
    ap.add_argument("--amg-rtol", type=float, default=0.15,
                    help="GMRES relative tolerance of the AMG form (on the 2M wing 0.15-0.2 reach the drop "
                         "in less time than 0.1 or 0.05, 0.4-0.5 in more; the ILU(0) forms are the same at "
                         "0.05 and 0.1)")
    ap.add_argument("--restart", type=int, default=30, help="GMRES restart length of every implicit form")
    ap.add_argument("--drop", type=float, default=1e-8, help="residual drop of the time-to-drop runs")
    ap.add_argument("--forms", default="ilu0,learned,learned_untrained,amg",
                    help="mixed forms of the time-to-drop runs besides the plain one (ilu0, learned, "
                         "learned_untrained, learned_handset, amg)")
    ap.add_argument("--lp-train-steps", type=int, default=2, help="Adam steps per system of the learned training run")
    ap.add_argument("--lp-train-pseudo", type=int, default=120,
                    help="pseudo-time steps (systems) of the learned training run")
    ap.add_argument("--lp-lr", type=float, default=3e-3, help="Adam learning rate of the learned training")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false", help="keep the default launch shapes")
    ap.add_argument("--rtol", type=float, default=5e-2,
                    help="GMRES relative tolerance of every form (the outer loop needs no tighter)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU baseline: timed steps continue until this much time has passed (at least one)")
    ap.add_argument("--max-ramp", type=int, default=300,
                    help="implicit: untimed ramp steps allowed before the Courant number reaches its cap")
    ap.add_argument("--config1-iters", type=int, default=2000,
                    help="iterations of the config-1 explicit GPU-vs-CPU comparison")
    ap.add_argument("--no-config1", action="store_true", help="skip the config-1 explicit comparison")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-drop", action="store_true")
    args = ap.parse_args()
    # the CPU baseline's OpenMP placement, read by the runtime when the oracle loads
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    if args.copy_probe:
        copy_probe()
        return
    implicit = args.workload.startswith("implicit")
    if args.steps is None:
        args.steps = 20
    if args.warmup is None:
        args.warmup = 5 if implicit else 3
    if args.gpus != 1:
        print(json.dumps({"error": "this solver core runs on one GPU; the multi-GPU (partitioned mesh, "
                          "halo exchange) form is designed but not built", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "implicit-sweep":
        out = implicit_sweep_workload(args)
    elif args.workload == "implicit-bump":
        out = implicit_bump_workload(args)
    elif args.workload.startswith("implicit"):
        out = implicit_workload(args)
    elif args.workload == "explicit-sweep":
        out = sweep_workload(args)
    else:
        out = explicit_workload(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
