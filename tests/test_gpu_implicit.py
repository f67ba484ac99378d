"""GPU parity of the implicit path (Jacobians, assembly, linear solve, update).

The assembled system is compared bitwise with the reference's own CSR matrix
(tests/golden, produced by the reference's MatrixSparse and Eigen AD) and with
the oracle.  The reference's AMGCL solve cannot be built here (Boost absent),
so the GMRES solution is pinned by a direct sparse solve of the same system
(scipy) and by its true residual.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from bes_amd import core, meshgen as mg
from conftest import golden_mesh
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _oracle_csr(om, U, cfl, finf):
    rp, col, val, rhs, _ = om.implicit_assemble(U, cfl, finf)
    nv = om.nv
    rows, cols, vals = [], [], []
    for e in range(om.ncell):
        for a in range(nv):
            for b in range(rp[e], rp[e + 1]):
                rows.extend([e * nv + a] * nv)
                cols.extend(col[b] * nv + np.arange(nv))
                vals.extend(val[b, a, :])
    A = sp.csr_matrix((vals, (rows, cols)), shape=(om.ncell * nv, om.ncell * nv))
    return A, rhs


def test_assembly_matches_reference(golden):
    name, g = golden
    print("case", name, flush=True)
    nd = int(g["ndim"])
    with core.Context(golden_mesh(g)) as c:
        c.set_freestream(orc.freestream_primitive(g["mach"], nd))
        c.set_state(g["state"])
        c.implicit_setup(64, "ilu0")
        lb, lr = c.implicit_assemble(float(g["cfl"]), local=True, repeat_nv=True)
        rp, col, val, rhs = c.implicit_system()
        assert np.array_equal(rp, g["csr_rowptr"])
        assert np.array_equal(col, g["csr_col"])
        assert np.array_equal(val, g["csr_val"])
        assert np.array_equal(rhs, g["impl_rhs"])
        assert np.array_equal(lb, np.abs(g["impl_rhs"]).max(0))
        assert np.array_equal(lr, np.abs(g["resid"]).max(0))


@pytest.mark.parametrize("mode", [orc.COMPAT, orc.EXACT])
def test_assembly_larger_matches_oracle(mode):
    md = mg.box_tets(10, 8, 6, seed=41, slip_faces=("zmin",))
    mach = (0.84, 0.0, 0.045)
    finf = orc.freestream_primitive(mach, 3)
    U = orc.perturbed_state(md.ncell, 3, mach, seed=42)
    om = orc.Mesh(md, mode)
    A, rhs = _oracle_csr(om, U, 25.0, finf)
    with core.Context(md, geom_mode=mode) as c:
        c.set_freestream(finf)
        c.set_state(U)
        c.implicit_setup(64, "jacobi")
        c.implicit_assemble(25.0)
        rp, col, val, rhs_g = c.implicit_system()
        Ag = sp.csr_matrix((val, col, rp), shape=A.shape)
        assert np.array_equal(rhs_g, rhs)
        assert abs(Ag - A).max() == 0


@pytest.mark.parametrize("precond", ["none", "jacobi", "ilu0"])
@pytest.mark.parametrize("which", ["tri", "box"])
def test_gmres_matches_direct_solve(which, precond):
    print("case", which, precond, flush=True)
    if which == "tri":
        md, mach = mg.bump_channel(30, 15, seed=3), (0.5, 0.0)
    else:
        md, mach = mg.box_tets(8, 6, 5, seed=43, slip_faces=("zmin",)), (0.84, 0.0, 0.045)
    nd = md.ndim
    finf = orc.freestream_primitive(mach, nd)
    U = orc.perturbed_state(md.ncell, nd, mach, seed=44)
    cfl = 20.0
    om = orc.Mesh(md)
    A, rhs = _oracle_csr(om, U, cfl, finf)
    xd = spla.spsolve(A.tocsc(), rhs.ravel()).reshape(rhs.shape)
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U)
        c.implicit_setup(64, precond, restart=30, maxiter=3000, rtol=1e-12)
        c.implicit_assemble(cfl)
        it, rr = c.linear_solve()
        x = c.linear_solution()
    assert rr <= 1e-12
    true = np.linalg.norm(rhs.ravel() - A @ x.ravel()) / np.linalg.norm(rhs)
    assert true <= 2e-12
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    print(which, precond, "iterations", it)


def test_preconditioners_reduce_iterations():
    md, mach = mg.box_tets(10, 8, 6, seed=45, slip_faces=("zmin",)), (0.84, 0.0, 0.045)
    finf = orc.freestream_primitive(mach, 3)
    U = orc.perturbed_state(md.ncell, 3, mach, seed=46)
    its = {}
    with core.Context(md) as c:
        c.set_freestream(finf)
        for pc in ("none", "jacobi", "ilu0"):
            c.set_state(U)
            c.implicit_setup(64, pc, rtol=1e-8, maxiter=3000)
            c.implicit_assemble(200.0)
            its[pc] = c.linear_solve()[0]
    print(its)
    assert its["ilu0"] <= its["jacobi"] <= its["none"]


def test_fp32_inner_solve():
    md, mach = mg.box_tets(8, 6, 5, seed=47, slip_faces=("zmin",)), (0.84, 0.0, 0.045)
    finf = orc.freestream_primitive(mach, 3)
    U = orc.perturbed_state(md.ncell, 3, mach, seed=48)
    om = orc.Mesh(md)
    A, rhs = _oracle_csr(om, U, 50.0, finf)
    xd = spla.spsolve(A.tocsc(), rhs.ravel()).reshape(rhs.shape)
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U)
        c.implicit_setup(32, "ilu0", rtol=1e-5)
        c.implicit_assemble(50.0)
        rp, col, val, rhs_g = c.implicit_system()
        Ag = sp.csr_matrix((val, col, rp), shape=A.shape)
        Ag.sort_indices()
        A2 = A.tocsr()
        A2.sort_indices()
        assert np.array_equal(Ag.indices, A2.indices) and np.array_equal(Ag.indptr, A2.indptr)
        # FP32 assembly: the face Jacobians are stored in FP32 before the block
        # sums and the blocks are stored in FP32, so each nv x nv block agrees
        # with the FP64 assembly to a few FP32 roundings of its largest entry
        # (a diagonal block sums the cell's face terms, which can cancel)
        nv = 5
        bg = Ag.toarray().reshape(-1, nv, A.shape[0] // nv, nv).transpose(0, 2, 1, 3)
        b64 = A2.toarray().reshape(-1, nv, A.shape[0] // nv, nv).transpose(0, 2, 1, 3)
        bmax = np.abs(b64).max(axis=(2, 3), keepdims=True)
        assert (np.abs(bg - b64) <= 3e-7 * bmax).all()
        assert (np.abs(bg - b64) > 0).any()    # really rounded
        assert np.array_equal(rhs_g, rhs)      # right-hand side stays FP64
        it, rr = c.linear_solve()
        x = c.linear_solution()
    assert rr <= 1e-5
    assert np.linalg.norm(x - xd) <= 1e-4 * np.linalg.norm(xd)


def test_implicit_step_matches_oracle():
    md, mach = mg.bump_channel(40, 20, seed=6), (0.5, 0.0)
    finf = orc.freestream_primitive(mach, 2)
    U = orc.perturbed_state(md.ncell, 2, mach, seed=49)
    om = orc.Mesh(md)
    A, rhs = _oracle_csr(om, U, 10.0, finf)
    Unew = U + spla.spsolve(A.tocsc(), rhs.ravel()).reshape(rhs.shape)
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U)
        c.implicit_setup(64, "ilu0", rtol=1e-13, maxiter=2000)
        c.implicit_assemble(10.0)
        c.linear_solve()
        c.implicit_update()
        Ug = c.get_state()
    assert np.abs(Ug - Unew).max() <= 1e-11 * np.abs(Unew).max()


@pytest.mark.parametrize("precision,repeat", [(64, True), (64, False), (32, False)])
def test_implicit_run_converges(precision, repeat):
    """The reference's nV-fold block repeat limits a step to 1/nV of a Newton
    step at large CFL, so it converges slowly; without it the residual drops fast."""
    md, mach = mg.bump_channel(48, 24, seed=7), (0.5, 0.0)
    finf = orc.freestream_primitive(mach, 2)
    U = np.tile(orc.conserved_from_primitive(finf, 2), (md.ncell, 1))
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U)
        c.implicit_setup(precision, "ilu0", rtol=1e-3, maxiter=300)
        hist = c.implicit_run(80, 10.0, 1000.0, (1.1, 0.75), stop_drop=1e-8, repeat_nv=repeat)
    nv = 4
    r = hist[:, nv]    # Linf of the density residual sum F*A
    print(precision, repeat, len(hist), r[0], r[-1], hist[-1, 2 * nv:])
    assert np.isfinite(hist).all()
    assert r[-1] <= (1e-3 if repeat else 1e-8) * r[0]
