"""The persistent explicit step (bes_explicit_persistent): every pass of every
pseudo-time step in one cooperative launch, the grid's blocks looping over the
passes' blocks and meeting at grid barriers.  Its states and Linf monitors must
be bitwise those of the launched passes (themselves the oracle's), for RK4 and
for a one-stage scheme over an odd number of steps (the state slots swap), on a
mesh small enough for one block per pass-block and on one with more blocks
than the grid holds, with and without a frozen limiter."""
import numpy as np
import pytest

from bes_amd import core, meshgen as mg
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _run(md, mach, U, niter, alpha, persistent, freeze=None):
    finf = orc.freestream_primitive(mach, md.ndim)
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U)
        if freeze is not None:
            c.freeze_limiter_values(freeze)
        c.explicit_persistent(persistent)
        linf = c.explicit_run(niter, 0.8, alpha=alpha)
        return c.get_state(), linf


@pytest.mark.parametrize("case", ["bump_rk4", "bump_euler3", "box_rk4", "box_big_rk4", "box_frozen"])
def test_persistent_matches_launched(case):
    print("case", case, flush=True)
    if case.startswith("bump"):
        md, mach = mg.bump_channel(40, 20, seed=5), (0.5, 0.0)
    elif case == "box_big_rk4":
        md, mach = mg.box_tets(40, 32, 32, seed=9, slip_faces=("zmin",)), (0.84, 0.0, 0.045)   # 245,760 tets
    else:
        md, mach = mg.box_tets(12, 10, 8, seed=61, slip_faces=("zmin",)), (0.84, 0.0, 0.045)
    alpha, niter = (core.RK4_ALPHA, 4) if "euler" not in case else ((1.0,), 3)
    U = orc.perturbed_state(md.ncell, md.ndim, mach, seed=17)
    freeze = None
    if case == "box_frozen":
        with core.Context(md) as c:
            c.set_freestream(orc.freestream_primitive(mach, md.ndim))
            c.set_state(U)
            _, freeze = c.gradient_limiter()
    Ul, ll = _run(md, mach, U, niter, alpha, False, freeze)
    Up, lp = _run(md, mach, U, niter, alpha, True, freeze)
    assert np.array_equal(Up, Ul) and np.array_equal(lp, ll)
    assert not np.array_equal(Up, U)
    if case == "bump_rk4":   # and the oracle's loop
        finf = orc.freestream_primitive(mach, 2)
        Uo, lo = orc.Mesh(md).explicit(U, niter, 0.8, finf, nthreads=1)
        assert np.array_equal(Up, Uo) and np.array_equal(lp, lo)
