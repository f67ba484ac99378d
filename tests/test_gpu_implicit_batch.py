"""The implicit path over a batch of freestream conditions (one launch per
kernel for all of them) against each condition run alone.

The reference runs one condition per process (Config.cpp:479-488) through the
loop of SolverImplicit.cpp:552-600; a batch must give every condition exactly
what its own run gives: the same face Jacobians, systems, factors, Krylov
iterates (per-condition Arnoldi processes, stopping points and restarts, and
reductions whose order does not depend on the batch size), Courant numbers and
states, bit for bit, including conditions that reach their drop early and are
frozen while the others go on.
"""
import numpy as np
import pytest

from bes_amd import core, meshgen as mg

pytestmark = pytest.mark.gpu

CONDS = [(0.5, 0.0), (0.84, 3.06), (0.7, -2.0), (0.78, 5.5)]   # Mach, angle of attack (deg)


def _freestream(mach, aoa):
    P = core.freestream_primitive(core.wing_freestream(mach, aoa), 3)
    return P, core.conserved_from_primitive(P, 3)


def _single(md, q, prec, precond, eta, nsteps, drop, restart):
    P, U = _freestream(*CONDS[q])
    with core.Context(md) as c:
        c.set_freestream(P)
        c.set_state(np.tile(U, (md.ncell, 1)))
        c.implicit_setup(prec, precond, restart=restart, maxiter=200, rtol=5e-2)
        c.implicit_set_reorth(eta)
        h = c.implicit_run(nsteps, 1.0, 1000.0, (0.5, 1.5), stop_drop=drop, repeat_nv=False)
        return h, c.get_state()


@pytest.mark.parametrize("prec,precond,eta,restart", [(64, "ilu0", 0.0, 30), (32, "ilu0", 0.0, 30),
                                                      (64, "jacobi", 0.7, 10)])
def test_batch_equals_single_runs(prec, precond, eta, restart):
    print("case", prec, precond, eta, restart, flush=True)
    md = mg.swept_wing(12, 6, 8)
    B, nsteps, drop = len(CONDS), 40, 1e-6
    with core.Context(md) as c:
        c.set_batch(B)
        P = np.stack([_freestream(*q)[0] for q in CONDS])
        c.set_freestream(P)
        c.set_state(np.stack([np.tile(_freestream(*q)[1], (md.ncell, 1)) for q in CONDS]))
        c.implicit_setup(prec, precond, restart=restart, maxiter=200, rtol=5e-2)
        c.implicit_set_reorth(eta)
        H = c.implicit_run(nsteps, 1.0, 1000.0, (0.5, 1.5), stop_drop=drop, repeat_nv=False)
        Ub = c.get_state()
        st = c.implicit_batch_state()
    assert H.shape[1:] == (B, 2 * 5 + 3)
    for q in range(B):
        h, U = _single(md, q, prec, precond, eta, nsteps, drop, restart)
        n = int(st["steps"][q])
        print(q, "steps", n, len(h), "reached", st["reached"][q], "krylov", int(np.nansum(H[:n, q, 11])), flush=True)
        assert n == len(h)
        assert np.array_equal(H[:n, q], h), f"condition {q}: history differs"
        assert np.isnan(H[n:, q]).all()
        assert np.array_equal(Ub[q], U), f"condition {q}: state differs"
    assert st["reached"].any()


def test_batch_assembly_and_solve_per_condition():
    """One assembly and one FP64 ILU(0)-GMRES solve of a batch of two different
    states: each condition's right-hand side and solution equal its own."""
    md = mg.swept_wing(12, 6, 8)
    rng = np.random.default_rng(5)
    Ps = [_freestream(*CONDS[1])[0], _freestream(*CONDS[3])[0]]
    Us = [np.tile(_freestream(*CONDS[q])[1], (md.ncell, 1)) * (1 + 0.01 * rng.uniform(-1, 1, (md.ncell, 5)))
          for q in (1, 3)]
    out = []
    for q in range(2):
        with core.Context(md) as c:
            c.set_freestream(Ps[q])
            c.set_state(Us[q])
            c.implicit_setup(64, "ilu0", rtol=1e-10, maxiter=500)
            lb, lr = c.implicit_assemble(50.0)
            it, rr = c.linear_solve()
            out.append((lb, lr, it, c.linear_solution()))
    with core.Context(md) as c:
        c.set_batch(2)
        c.set_freestream(np.stack(Ps))
        c.set_state(np.stack(Us))
        c.implicit_setup(64, "ilu0", rtol=1e-10, maxiter=500)
        lb, lr = c.implicit_assemble(50.0)
        it, rr = c.linear_solve()
        x = c.linear_solution()
        st = c.implicit_batch_state()
    for q in range(2):
        assert np.array_equal(lb[q], out[q][0]) and np.array_equal(lr[q], out[q][1])
        assert st["iters"][q] == out[q][2]
        assert np.array_equal(x[q], out[q][3])
    assert it == max(o[2] for o in out)


def test_batch_rejects_what_it_does_not_cover():
    md = mg.swept_wing(8, 4, 6)
    with core.Context(md) as c:
        c.set_batch(2)
        with pytest.raises(core.BesError):
            c.implicit_setup(32, "amg")
        c.implicit_setup(64, "ilu0")
        with pytest.raises(core.BesError):
            c.implicit_set_solver("bicgstab")
