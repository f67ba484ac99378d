"""Krylov solvers of CoefficientMatrix::solve (CoefficientMatrix.cpp:210-325) on the GPU.

The reference solves with AMGCL, which cannot be built here (it needs Boost),
so every solve is pinned by a direct sparse solve of the same system (scipy)
and by its true residual; the GMRES iterates are also checked against a plain
numpy restatement of restarted right-preconditioned GMRES when the iteration
limit ends a solve in the middle of a cycle.  Both orthogonalisation modes are
covered: classical Gram-Schmidt once, with the next Arnoldi step queued before
the host reads the Hessenberg column (eta 0), and with DGKS re-orthogonalisation
(eta 0.7).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from bes_amd import core, meshgen as mg
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _system(seed=43, cfl=20.0):
    md, mach = mg.box_tets(8, 6, 5, seed=seed, slip_faces=("zmin",)), (0.84, 0.0, 0.045)
    finf = orc.freestream_primitive(mach, 3)
    U = orc.perturbed_state(md.ncell, 3, mach, seed=seed + 1)
    om = orc.Mesh(md)
    rp, col, val, rhs, _ = om.implicit_assemble(U, cfl, finf)
    A = sp.bsr_matrix((val, col, rp), shape=(md.ncell * 5,) * 2).tocsr()
    return md, finf, U, A, rhs


def _solve(md, finf, U, cfl, precision, precond, solver, eta, restart=30, maxiter=3000, rtol=1e-12,
           allow_unconverged=False):
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U)
        c.implicit_setup(precision, precond, restart=restart, maxiter=maxiter, rtol=rtol)
        c.implicit_set_reorth(eta)
        c.implicit_set_solver(solver)
        c.implicit_assemble(cfl)
        it, rr = c.linear_solve(allow_unconverged=allow_unconverged)
        x = c.linear_solution()
        st = c.implicit_stats()
    return it, rr, x, st


@pytest.mark.parametrize("eta", [0.0, 0.7])
@pytest.mark.parametrize("precond", ["none", "jacobi", "ilu0"])
@pytest.mark.parametrize("solver", ["gmres", "bicgstab"])
def test_krylov_fp64_matches_direct_solve(solver, precond, eta):
    if solver == "bicgstab" and eta > 0:
        pytest.skip("the re-orthogonalisation threshold applies to GMRES only")
    print("case", solver, precond, eta, flush=True)
    md, finf, U, A, rhs = _system()
    xd = spla.spsolve(A.tocsc(), rhs.ravel()).reshape(rhs.shape)
    rtol = 1e-12 if solver == "gmres" else 1e-11
    it, rr, x, st = _solve(md, finf, U, 20.0, 64, precond, solver, eta, rtol=rtol)
    true = np.linalg.norm(rhs.ravel() - A @ x.ravel()) / np.linalg.norm(rhs)
    print(solver, precond, eta, "iterations", it, "relres", rr, "true", true, st)
    assert rr <= rtol * 1.0001
    np.testing.assert_allclose(true, rr, rtol=1e-3, atol=1e-15)
    assert np.linalg.norm(x - xd) <= 1e4 * rtol * np.linalg.norm(xd)


@pytest.mark.parametrize("eta", [0.0, 0.7])
@pytest.mark.parametrize("solver", ["gmres", "bicgstab"])
def test_krylov_fp32_inner_solve(solver, eta):
    if solver == "bicgstab" and eta > 0:
        pytest.skip("the re-orthogonalisation threshold applies to GMRES only")
    md, finf, U, A, rhs = _system(seed=47, cfl=50.0)
    xd = spla.spsolve(A.tocsc(), rhs.ravel()).reshape(rhs.shape)
    it, rr, x, st = _solve(md, finf, U, 50.0, 32, "ilu0", solver, eta, rtol=1e-5)
    true = np.linalg.norm(rhs.ravel() - A @ x.ravel()) / np.linalg.norm(rhs)
    print(solver, eta, it, rr, true, st)
    assert rr <= 1e-5 * 1.0001
    assert true <= 1.01e-5      # the FP32 matrix differs from A by rounding only
    assert np.linalg.norm(x - xd) <= 1e-4 * np.linalg.norm(xd)


@pytest.mark.parametrize("eta", [0.0, 0.7])
def test_gmres_several_cycles(eta):
    """restart 5 with ~40 iterations: every cycle restarts from the true residual."""
    md, finf, U, A, rhs = _system(seed=51)
    xd = spla.spsolve(A.tocsc(), rhs.ravel()).reshape(rhs.shape)
    it, rr, x, st = _solve(md, finf, U, 20.0, 64, "jacobi", "gmres", eta, restart=5, rtol=1e-10)
    true = np.linalg.norm(rhs.ravel() - A @ x.ravel()) / np.linalg.norm(rhs)
    print(eta, it, rr, true, st)
    assert it > 10
    assert rr <= 1e-10 * 1.0001
    np.testing.assert_allclose(true, rr, rtol=1e-3)
    assert np.linalg.norm(x - xd) <= 1e-6 * np.linalg.norm(xd)


def _np_gmres(A, b, m, maxiter):
    """Restarted GMRES(m), no preconditioner, x0 = 0: the least-squares problem of
    each cycle solved directly (the device rotates it with Givens rotations)."""
    x = np.zeros_like(b)
    total = 0
    while total < maxiter:
        r = b - A @ x
        beta = np.linalg.norm(r)
        V = [r / beta]
        H = np.zeros((m + 1, m))
        j = 0
        while j < m and total < maxiter:
            w = A @ V[j]
            h = np.array([v @ w for v in V])
            w = w - sum(hq * v for hq, v in zip(h, V))
            H[:j + 1, j] = h
            H[j + 1, j] = np.linalg.norm(w)
            V.append(w / H[j + 1, j])
            total += 1
            j += 1
        e1 = np.zeros(j + 1)
        e1[0] = beta
        y = np.linalg.lstsq(H[:j + 1, :j], e1, rcond=None)[0]
        x = x + np.array(V[:j]).T @ y
    return x


@pytest.mark.parametrize("eta", [0.0, 0.7])
@pytest.mark.parametrize("maxiter", [23, 10, 1])
def test_gmres_stops_at_maxiter_mid_cycle(eta, maxiter):
    """The iteration limit ends the solve inside a cycle (23 with restart 10),
    at a cycle end (10) and after one step (1): the device discards the Arnoldi
    step it queued ahead, and its x equals restarted GMRES's iterate."""
    md, finf, U, A, rhs = _system(seed=55, cfl=200.0)
    it, rr, x, st = _solve(md, finf, U, 200.0, 64, "none", "gmres", eta, restart=10, maxiter=maxiter,
                           rtol=1e-14, allow_unconverged=True)
    b = rhs.ravel()
    xr = _np_gmres(A, b, 10, maxiter)
    true = np.linalg.norm(b - A @ x.ravel()) / np.linalg.norm(b)
    print(eta, maxiter, it, rr, true, np.linalg.norm(x.ravel() - xr) / np.linalg.norm(xr))
    assert it == maxiter
    assert 0 < rr < 1
    np.testing.assert_allclose(true, rr, rtol=1e-6)
    assert np.linalg.norm(x.ravel() - xr) <= 1e-8 * np.linalg.norm(xr)


def test_lgmres_is_rejected():
    md, finf, U, A, rhs = _system()
    with core.Context(md) as c:
        c.implicit_setup(64, "ilu0")
        with pytest.raises(core.BesError) as e:
            lib = core.lib()
            core._check(lib.bes_implicit_set_solver(c.h, 2))
        assert "LGMRES" in str(e.value)
