"""Aggregation AMG preconditioner on the GPU, through the C ABI.

The GPU chooses its aggregates in parallel rounds, so they are checked by the
properties that define them (kept rows are exactly those with a strong
coupling; roots at least three strong links apart; every member within two
links of its root; every root's strong neighbours in its aggregate).  Every
other part of the hierarchy is compared with the CPU restatement
(oracle/amg.py) built from the GPU's aggregate ids: the Galerkin operators of
every level, and one V-cycle for each relaxation.  Solves are compared with a
direct solve, and iteration counts with the restatement's row-sweep
aggregation.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from bes_amd import core, meshgen as mg
from oracle import amg, oracle as orc

pytestmark = pytest.mark.gpu

MVEC = (0.84 * np.cos(np.radians(3.06)), 0.0, 0.84 * np.sin(np.radians(3.06)))


def _context(dims, precision=64, cfl=20.0, relax="jacobi", seed=3, rtol=1e-8, maxiter=400, repeat=True):
    md = mg.swept_wing(*dims)
    finf = orc.freestream_primitive(MVEC, 3)
    U = orc.perturbed_state(md.ncell, 3, MVEC, seed=seed, amp=0.02)
    c = core.Context(md)
    c.set_freestream(finf)
    c.implicit_setup(precision, "amg", restart=30, maxiter=maxiter, rtol=rtol)
    c.amg_set_params(relax=relax)
    c.set_state(U)
    c.implicit_assemble(cfl, repeat_nv=repeat)
    rp, col, val, rhs = c.implicit_system()
    A = amg.BlockCsr.from_scalar_csr(rp, col, val, c.nv)
    return c, A, rhs.ravel()


def _sym_graph(A, keep):
    st = amg.strong_connections(A)
    r = A.rows()
    adj = [set() for _ in range(A.n)]
    for p in np.flatnonzero(st):
        i, j = r[p], A.col[p]
        if keep[i] and keep[j]:
            adj[i].add(j)
            adj[j].add(i)
    return st, adj


def _check_aggregates(A, agg, state):
    st = amg.strong_connections(A)
    r = A.rows()
    has_strong = np.zeros(A.n, bool)
    has_strong[r[st]] = True
    assert np.array_equal(state == 3, ~has_strong)
    assert np.array_equal(agg < 0, ~has_strong)
    _, adj = _sym_graph(A, has_strong)
    roots = np.flatnonzero(state == 1)
    nagg = agg.max() + 1
    assert len(roots) == nagg and len(np.unique(agg[roots])) == nagg
    root_of = np.full(nagg, -1)
    root_of[agg[roots]] = roots
    for q in roots:
        assert all(agg[j] == agg[q] for j in adj[q])          # a root holds its strong neighbours
        two = set().union(*[adj[j] for j in adj[q]]) if adj[q] else set()
        assert not any(state[j] == 1 for j in (adj[q] | two) - {q})   # roots three links apart
    for i in np.flatnonzero(state == 2):
        q = root_of[agg[i]]
        assert i in adj[q] or any(j in adj[q] and agg[j] == agg[i] for j in adj[i])
    return nagg


def _oracle_from_gpu(c, A, kind):
    rows, nnzb, direct = c.amg_info()
    aggs = [c.amg_level(l)[0] for l in range(len(rows))]
    levels, inv = amg.hierarchy(A, aggregates=lambda l, Al: (int(rows[l + 1]), aggs[l]), kind=kind)
    return rows, direct, levels, inv


@pytest.mark.parametrize("dims", [(8, 4, 6), (16, 8, 12)])
def test_hierarchy_matches_restatement(dims):
    c, A, b = _context(dims)
    with c:
        z = c.amg_apply(b)
        rows, nnzb, direct = c.amg_info()
        print("levels", list(rows), "blocks", list(nnzb), "direct", direct)
        assert rows[0] == A.n and len(rows) >= 2 and direct
        assert rows[-1] * c.nv <= 500 < rows[-2] * c.nv
        # level 0: aggregates from the strength of the assembled matrix
        agg0, st0, _ = c.amg_level(0)
        assert _check_aggregates(A, agg0, st0) == rows[1]
        # every coarse operator against the block sums of the level above
        prev, prev_agg = A, agg0
        for l in range(1, len(rows)):
            aggl, stl, (rp, col, val) = c.amg_level(l)
            ref = amg.galerkin(prev, prev_agg, int(rows[l]))
            assert np.array_equal(rp, ref.rp) and np.array_equal(col, ref.col)
            assert np.abs(val - ref.val).max() <= 1e-12 * np.abs(ref.val).max()
            Al = amg.BlockCsr(rp, col, val)
            if l < len(rows) - 1:
                assert _check_aggregates(Al, aggl, stl) == rows[l + 1]
            prev, prev_agg = Al, aggl
        # one V-cycle against the restatement built from the same aggregates
        _, _, levels, inv = _oracle_from_gpu(c, A, "jacobi")
        assert len(levels) == len(rows)
        zr = amg.vcycle(levels, inv, b).reshape(z.shape)
        err = np.abs(z - zr).max() / np.abs(zr).max()
        print("V-cycle rel err %.2e" % err)
        assert err <= 1e-9


@pytest.mark.parametrize("relax,kind", [("jacobi", "jacobi"), ("block_jacobi", "block")])
def test_row_sweep_hierarchy_is_amgcl(relax, kind):
    """With AMGCL's row sweep the GPU builds the reference's hierarchy itself:
    the aggregates of every level equal the restatement's (in the reference's
    cell order), and a V-cycle matches the restatement's own hierarchy."""
    c, A, b = _context((16, 8, 12), relax=relax)
    with c:
        c.amg_set_aggregation("row_sweep")
        z = c.amg_apply(b).ravel()
        rows, nnzb, direct = c.amg_info()
        levels, inv = amg.hierarchy(A, kind=kind)
        assert len(levels) == len(rows) and (inv is not None) == direct
        for l in range(len(rows) - 1):
            agg = c.amg_level(l)[0]
            assert np.array_equal(agg, levels[l].agg), f"level {l} aggregates"
        zr = amg.vcycle(levels, inv, b)
        err = np.abs(z - zr).max() / np.abs(zr).max()
        print("levels", list(rows), "V-cycle rel err %.2e" % err)
        assert err <= 1e-9


@pytest.mark.parametrize("relax,kind", [("spai0", "spai0"), ("block_jacobi", "block"), ("jacobi", "jacobi")])
def test_cycle_relaxations_match_restatement(relax, kind):
    c, A, b = _context((16, 8, 12), relax=relax)
    with c:
        rng = np.random.default_rng(8)
        r = rng.standard_normal(b.shape)
        z = c.amg_apply(r).ravel()
        _, _, levels, inv = _oracle_from_gpu(c, A, kind)
        zr = amg.vcycle(levels, inv, r)
        assert np.abs(z - zr).max() <= 1e-9 * np.abs(zr).max()


def test_fp32_cycle_close_to_restatement():
    """FP32 solves run the level-0 sweeps from row-scaled FP16 copies of the
    matrix and relaxation blocks; the cycle stays within that rounding of the
    FP64 restatement built from the same aggregates."""
    c, A, b = _context((16, 8, 12), precision=32, relax="block_jacobi")
    with c:
        rng = np.random.default_rng(9)
        r = rng.standard_normal(b.shape)
        z = c.amg_apply(r).ravel()
        _, _, levels, inv = _oracle_from_gpu(c, A, "block")
        zr = amg.vcycle(levels, inv, r)
        rel = np.linalg.norm(z - zr) / np.linalg.norm(zr)
        print("FP32 cycle rel err %.2e" % rel)
        assert rel < 1e-2


def test_gmres_amg_solves_and_compares_with_row_sweep():
    c, A, b = _context((16, 8, 12), cfl=20.0, relax="block_jacobi", rtol=1e-8, repeat=False)
    with c:
        it, rr = c.linear_solve()
        x = c.linear_solution().ravel()
    S = A.scalar().tocsc()
    xd = spla.spsolve(S, b)
    assert rr <= 1e-8 * 1.0001
    assert np.abs(x - xd).max() <= 1e-6 * np.abs(xd).max()
    levels, inv = amg.hierarchy(A, kind="block")
    _, it_ref, _ = amg.gmres(S, lambda r: amg.vcycle(levels, inv, r), b, rtol=1e-8, maxiter=400)
    print(f"GMRES(30)+AMG iterations: GPU (parallel aggregates) {it}, row-sweep aggregates {it_ref}")
    assert it <= 1.5 * it_ref + 5


def test_fp32_amg_solve():
    c, A, b = _context((16, 8, 12), precision=32, cfl=20.0, relax="block_jacobi", rtol=1e-6, repeat=False)
    with c:
        it, rr = c.linear_solve()
        x = c.linear_solution().ravel()
    xd = spla.spsolve(A.scalar().tocsc(), b)
    assert rr <= 1e-6 * 1.0001
    assert np.abs(x - xd).max() <= 1e-4 * np.abs(xd).max()


def test_lagged_hierarchy():
    """rebuild_interval 3: three assemble-and-solve rounds build one hierarchy,
    every solve still reaches its tolerance, and new parameters force a build."""
    md = mg.bump_channel(48, 24, seed=7)
    finf = orc.freestream_primitive((0.5, 0.0), 2)
    U = orc.perturbed_state(md.ncell, 2, (0.5, 0.0), seed=2, amp=0.02)
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U)
        c.implicit_setup(32, "amg", rtol=1e-4, maxiter=300)
        c.amg_set_params(relax="block_jacobi", rebuild_interval=3)
        its = []
        for cfl in (10.0, 12.0, 14.0, 16.0):
            c.implicit_assemble(cfl, repeat_nv=False)
            it, rr = c.linear_solve()
            assert rr <= 1e-4 * 1.0001
            its.append(it)
            c.implicit_update()
        assert c.amg_info(with_builds=True)[3] == 2   # solves 1 and 4
        c.amg_set_params(relax="block_jacobi", rebuild_interval=3)
        c.implicit_assemble(16.0, repeat_nv=False)
        c.linear_solve()
        assert c.amg_info(with_builds=True)[3] == 3
        print("iterations with a lagged hierarchy", its)


def test_amg_implicit_run_converges():
    md, mach = mg.bump_channel(48, 24, seed=7), (0.5, 0.0)
    finf = orc.freestream_primitive(mach, 2)
    U = np.tile(orc.conserved_from_primitive(finf, 2), (md.ncell, 1))
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U)
        c.implicit_setup(64, "amg", rtol=1e-3, maxiter=300)
        c.amg_set_params(relax="block_jacobi")
        hist = c.implicit_run(80, 10.0, 1000.0, (1.1, 0.75), stop_drop=1e-8, repeat_nv=False)
    r = hist[:, 4]
    print(len(hist), r[0], r[-1])
    assert np.isfinite(hist).all() and r[-1] <= 1e-8 * r[0]


def test_reference_relaxation_at_low_cfl():
    """The reference's configuration (scalar damped Jacobi 0.72) converges at
    CFL 1 with its nV-fold block repeat, the start of its test runs; at larger
    CFL numbers (or on the next larger mesh) it stalls, with either aggregation."""
    c, A, b = _context((8, 4, 6), cfl=1.0, relax="jacobi", rtol=1e-6, repeat=True)
    with c:
        it, rr = c.linear_solve()
    levels, inv = amg.hierarchy(A)
    _, it_ref, rr_ref = amg.gmres(A.scalar(), lambda r: amg.vcycle(levels, inv, r), b, rtol=1e-6, maxiter=400)
    print(f"damped Jacobi AMG at CFL 1: GPU {it} iterations, row sweep {it_ref}")
    assert rr <= 1e-6 * 1.0001 and rr_ref <= 1e-6


def test_amg_arguments():
    md = mg.swept_wing(8, 4, 6)
    with core.Context(md) as c:
        c.set_freestream(orc.freestream_primitive(MVEC, 3))
        with pytest.raises(core.BesError):
            c.amg_set_params()          # before implicit_setup
        c.implicit_setup(64, "amg")
        for bad in (dict(eps_strong=0.0), dict(over_interp=0.5), dict(damping=0.0), dict(coarse_enough=0)):
            with pytest.raises(core.BesError):
                c.amg_set_params(**bad)
        with pytest.raises(core.BesError):
            c.amg_info()                # nothing built yet
