"""Limiter freeze (bes_set_limiter_freeze; no reference counterpart).

Once a condition's density residual drop reaches the freeze level in the
implicit loop (SolverImplicit.cpp:552-600), its Venkatakrishnan limiter
(LimiterVenkatakrishnan.cpp:21-99) keeps the values of that step's gradient
pass while the gradient is still recomputed.  Off (the default) nothing
changes; a batch freezes each condition at its own step, with the same bits
as the condition's own run.
"""
import numpy as np
import pytest

from bes_amd import core, meshgen as mg

pytestmark = pytest.mark.gpu

SCHEDULE = (1.0, 1000.0, (0.5, 1.5))
CONDS = [(0.84, 3.06), (0.6, -1.0)]


def _start(md, q=0):
    P = core.freestream_primitive(core.wing_freestream(*CONDS[q]), 3)
    return P, np.tile(core.conserved_from_primitive(P, 3), (md.ncell, 1))


def test_freeze_keeps_limiter_and_recomputes_gradient():
    md = mg.swept_wing(12, 6, 8)
    P, U0 = _start(md)
    with core.Context(md) as c:
        c.set_freestream(P)
        c.set_state(U0)
        c.implicit_setup(64, "ilu0", rtol=5e-2, maxiter=200)
        c.set_limiter_freeze(0.3)
        assert not c.limiter_frozen().any()
        h = c.implicit_run(40, *SCHEDULE, repeat_nv=False)
        drops = h[:, 5] / h[0, 5]
        assert drops.min() <= 0.3
        assert c.limiter_frozen().all()
        U = c.get_state()
        G1, P1 = c.gradient_limiter()
        rng = np.random.default_rng(7)
        c.set_state(U * (1 + 1e-3 * rng.uniform(-1, 1, U.shape)))
        G2, P2 = c.gradient_limiter()
        assert np.array_equal(P1, P2), "a frozen limiter changed"
        assert not np.array_equal(G1, G2)
        c.set_limiter_freeze(0.0)   # unfreeze: the limiter of the current state again
        G3, P3 = c.gradient_limiter()
        assert np.array_equal(G2, G3)
        Us = c.get_state()
    with core.Context(md) as c:
        c.set_freestream(P)
        c.set_state(Us)
        G4, P4 = c.gradient_limiter()
    assert np.array_equal(P3, P4) and not np.array_equal(P3, P1)


def test_freeze_off_changes_nothing():
    md = mg.swept_wing(12, 6, 8)
    P, U0 = _start(md)
    out = []
    for call in (False, True):
        with core.Context(md) as c:
            c.set_freestream(P)
            c.set_state(U0)
            c.implicit_setup(64, "ilu0", rtol=5e-2, maxiter=200)
            if call:
                c.set_limiter_freeze(0.0)
            out.append((c.implicit_run(15, *SCHEDULE, repeat_nv=False), c.get_state()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_freeze_in_a_batch_equals_single_runs():
    md = mg.swept_wing(12, 6, 8)
    nsteps, freeze = 30, 0.2
    singles = []
    for q in range(2):
        P, U0 = _start(md, q)
        with core.Context(md) as c:
            c.set_freestream(P)
            c.set_state(U0)
            c.implicit_setup(64, "ilu0", rtol=5e-2, maxiter=200)
            c.set_limiter_freeze(freeze)
            h = c.implicit_run(nsteps, *SCHEDULE, repeat_nv=False)
            singles.append((h, c.get_state(), bool(c.limiter_frozen()[0])))
    with core.Context(md) as c:
        c.set_batch(2)
        c.set_freestream(np.stack([_start(md, q)[0] for q in range(2)]))
        c.set_state(np.stack([_start(md, q)[1] for q in range(2)]))
        c.implicit_setup(64, "ilu0", rtol=5e-2, maxiter=200)
        c.set_limiter_freeze(freeze)
        H = c.implicit_run(nsteps, *SCHEDULE, repeat_nv=False)
        Ub = c.get_state()
        fr = c.limiter_frozen()
    for q in range(2):
        h, U, f = singles[q]
        print("condition", q, "frozen", f, "drop", h[-1, 5] / h[0, 5], flush=True)
        assert fr[q] == f
        assert np.array_equal(H[:len(h), q], h)
        assert np.array_equal(Ub[q], U)
    assert fr.any()


def test_freeze_argument_checks():
    md = mg.swept_wing(8, 4, 6)
    with core.Context(md) as c:
        for bad in (-0.1, 1.0, float("nan")):
            with pytest.raises(core.BesError):
                c.set_limiter_freeze(bad)
        c.set_limiter_freeze(0.5)
        c.set_batch(2)   # a batch change clears the freeze
        assert not c.limiter_frozen().any()


def test_freeze_at_given_values():
    """bes_freeze_limiter_values: every later gradient pass uses the given
    limiter (so several solver forms converge one frozen discretisation)."""
    md = mg.swept_wing(12, 6, 8)
    P, U0 = _start(md)
    rng = np.random.default_rng(11)
    with core.Context(md) as c:
        c.set_freestream(P)
        c.set_state(U0 * (1 + 1e-2 * rng.uniform(-1, 1, U0.shape)))
        G0, P0 = c.gradient_limiter()
        c.set_state(U0 * (1 + 1e-2 * rng.uniform(-1, 1, U0.shape)))
        G1, P1 = c.gradient_limiter()
        assert not np.array_equal(P0, P1)
        c.freeze_limiter_values(P0)
        assert c.limiter_frozen().all()
        G2, P2 = c.gradient_limiter()
        assert np.array_equal(P2, P0) and np.array_equal(G2, G1)
        c.implicit_setup(64, "ilu0", rtol=5e-2, maxiter=200)
        c.implicit_run(3, *SCHEDULE, repeat_nv=False)
        assert np.array_equal(c.gradient_limiter()[1], P0)
        c.set_limiter_freeze(0.0)
        assert not c.limiter_frozen().any()
        U = c.get_state()
        P3 = c.gradient_limiter()[1]
    with core.Context(md) as c:
        c.set_freestream(P)
        c.set_state(U)
        assert np.array_equal(c.gradient_limiter()[1], P3)
