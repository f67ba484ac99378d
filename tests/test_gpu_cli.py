"""The bes_hip driver (reference-style CLI over the C++ classes) end to end.

Writes a reference-format .besi and MSH2 mesh, runs ``bes_amd/bes_hip`` and
compares residual.csv (the reference's format, SolverExplicit.cpp:380-390,
SolverImplicit.cpp:328-343) and the solution files written after the first and
the last step (SolverExplicit.cpp:425-433, Solution.cpp:129-161) with the
oracle run on the same mesh.
"""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from bes_amd import core, meshgen as mg
from oracle import oracle as orc

from test_gpu_output import dimensional, read_csv, read_vtk

pytestmark = pytest.mark.gpu
CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bes_amd", "bes_hip")


def _run(tmp_path, md, solver, nsteps, extra=()):
    mg.write_msh2(str(tmp_path / "m.msh"), md)
    mg.write_besi(str(tmp_path / "c.besi"), "m.msh", md.ndim, mach=(0.5, 0.05, 0.0), solver=solver,
                  cfl=0.8 if solver == "EXPLICIT" else 5.0, cfl_max=0.0 if solver == "EXPLICIT" else 50.0,
                  exponents=(1.1, 0.75), nsteps=nsteps, output=".")
    if extra:
        with open(tmp_path / "c.besi", "a") as fh:
            fh.write("\n".join(extra) + "\n")
    res = subprocess.run([CLI, "c.besi"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    return res


def _csv(path):
    rows = [l.rstrip(",\n").split(",") for l in open(path) if l.strip()]
    return np.array([[float(x) for x in r] for r in rows])


def test_cli_explicit_matches_oracle(tmp_path):
    md = mg.bump_channel(30, 15, seed=9)
    res = _run(tmp_path, md, "EXPLICIT", 6)
    assert res.returncode == 0, res.stderr
    got = _csv(tmp_path / "residual.csv")
    finf = orc.freestream_primitive((0.5, 0.05), 2)
    U = np.tile(orc.conserved_from_primitive(finf, 2), (md.ncell, 1))
    Ufin, linf = orc.Mesh(md).explicit(U, 6, 0.8, finf, nthreads=1)
    assert got.shape == (6, 5)
    assert np.array_equal(got[:, 0], np.arange(1, 7))
    np.testing.assert_allclose(got[:, 1:], linf, rtol=6e-6, atol=1e-300)
    # SAVE_INTERVAL larger than the run: files after the first and the last step
    assert sorted(p.name for p in tmp_path.glob("solution_*")) == ["solution_1.vtk", "solution_6.vtk"]
    v = read_vtk(tmp_path / "solution_6.vtk")
    P = dimensional(Ufin, 2)
    np.testing.assert_allclose(v["rho_nondim"], P[:, 0], rtol=6e-6)
    np.testing.assert_allclose(v["P_nondim"], P[:, 3], rtol=6e-6)
    np.testing.assert_allclose(v["M"][:, :2], P[:, 1:3], rtol=6e-6, atol=1e-6)


@pytest.mark.parametrize("precond", [[], ["PRECONDITIONER = AMG", "AMG_PRECOND_COARSENER = AGGREGATION",
                                            "AMG_PRECOND_SMOOTHER = BLOCK_JACOBI"],
                                     ["LINEAR_SOLVER = BICGSTAB"], ["LINEAR_SOLVER = GMRES"]])
def test_cli_implicit_matches_oracle(tmp_path, precond):
    """Each step's solve against a direct solve of the oracle's system; the
    reference's Krylov choices (LINEAR_SOLVER, Config.cpp:345-358) included."""
    md = mg.bump_channel(24, 12, seed=10)
    res = _run(tmp_path, md, "IMPLICIT", 5, extra=["LINEAR_SOLVER_ABS_TOL = 1E-12",
                                                   "LINEAR_SOLVER_MAX_ITER = 2000",
                                                   "OUTPUT_FORMAT = CSV", "OUTPUT_SLIP_WALL = YES",
                                                   "SAVE_INTERVAL = 2"] + precond)
    assert res.returncode == 0, res.stderr
    got = _csv(tmp_path / "residual.csv")
    # oracle: reference implicit loop with a direct solve and the adaptive CFL
    finf = orc.freestream_primitive((0.5, 0.05), 2)
    om = orc.Mesh(md)
    U = np.tile(orc.conserved_from_primitive(finf, 2), (md.ncell, 1))
    cfl, old, ref = 5.0, np.zeros(4), []
    for _ in range(5):
        rp, col, val, rhs, _ = om.implicit_assemble(U, cfl, finf)
        A = sp.bsr_matrix((val, col, rp), shape=(md.ncell * 4,) * 2).tocsc()
        U = U + spla.spsolve(A, rhs.ravel()).reshape(U.shape)
        lb = np.abs(rhs).max(0)
        ref.append(lb)
        e = 0.75 if old[0] > lb[0] else 1.1   # exponent[1] when the residual fell (:537)
        cfl = min(50.0, max(5.0, cfl * (old[0] / lb[0]) ** e))
        old = lb
    assert got.shape == (5, 1 + 4 + 2)
    np.testing.assert_allclose(got[:, 1:5], np.array(ref), rtol=1e-5)
    # column 6: the true relative residual of each solve (BiCGStab stops on its
    # recursive residual, as AMGCL's does, so its true residual may sit a little above)
    assert (got[:, 6] <= (1e-11 if "BICGSTAB" in " ".join(precond) else 1e-12 * 1.0001)).all()
    # SAVE_INTERVAL 2: after steps 1, 3 and 5 (0-based 0, 2, 4), CSV and slip wall
    names = sorted(p.name for p in tmp_path.glob("s*_*.csv"))
    assert names == ["slip_wall_1.csv", "slip_wall_3.csv", "slip_wall_5.csv",
                     "solution_1.csv", "solution_3.csv", "solution_5.csv"]
    sol = read_csv(tmp_path / "solution_5.csv")
    np.testing.assert_allclose(sol[:, 2:], dimensional(U, 2), rtol=1e-5, atol=1e-7)
    wall = read_csv(tmp_path / "slip_wall_5.csv")
    assert wall.shape == ((om.faces()["btype"] == 2).sum(), 2 + 2 + 5)


def test_cli_rejects_unavailable_amg_smoother(tmp_path):
    md = mg.tri_patch(3, 3)
    res = _run(tmp_path, md, "IMPLICIT", 1, extra=["PRECONDITIONER = AMG", "AMG_PRECOND_SMOOTHER = GAUSS_SEIDEL"])
    assert res.returncode == 1 and "AMG_PRECOND_SMOOTHER" in res.stderr + res.stdout


def test_cli_rejects_lgmres(tmp_path):
    md = mg.tri_patch(3, 3)
    res = _run(tmp_path, md, "IMPLICIT", 1, extra=["LINEAR_SOLVER = LGMRES"])
    assert res.returncode == 1 and "LINEAR_SOLVER" in res.stderr + res.stdout


def test_cli_rejects_unknown_key(tmp_path):
    md = mg.tri_patch(3, 3)
    res = _run(tmp_path, md, "EXPLICIT", 1, extra=["NOT_A_KEY = 3"])
    assert res.returncode == 1
    assert "unknown config key" in res.stderr


def test_cli_limiter_freeze_key(tmp_path):
    """LIMITER_FREEZE_DROP (a key of this build) reaches bes_set_limiter_freeze:
    the driver's residual history is the Python binding's with the same freeze."""
    md = mg.bump_channel(24, 12, seed=10)
    n, drop = 8, 0.5
    res = _run(tmp_path, md, "IMPLICIT", n, extra=["LINEAR_SOLVER_ABS_TOL = 1E-10", "LINEAR_SOLVER_MAX_ITER = 2000",
                                                   f"LIMITER_FREEZE_DROP = {drop}"])
    assert res.returncode == 0, res.stderr
    got = _csv(tmp_path / "residual.csv")
    finf = orc.freestream_primitive((0.5, 0.05), 2)
    U0 = np.tile(orc.conserved_from_primitive(finf, 2), (md.ncell, 1))
    hist = {}
    for fr in (drop, 0.0):
        with core.Context(md) as c:
            c.set_freestream(finf)
            c.set_state(U0)
            c.implicit_setup(64, "ilu0", maxiter=2000, rtol=1e-10)
            c.set_limiter_freeze(fr)
            hist[fr] = c.implicit_run(n, 5.0, 50.0, (1.1, 0.75), repeat_nv=True)
            if fr > 0:
                assert c.limiter_frozen().all()
    np.testing.assert_allclose(got[:, 1:5], hist[drop][:, 0:4], rtol=1e-6)
    assert not np.array_equal(hist[drop], hist[0.0])
