"""The oracle's explicit loop against the reference's own units, run as a loop.

oracle/_ref/bes_ref_time drives the reference's Config, Mesh, SolutionSnapshot,
GradientGreenGauss, LimiterVenkatakrishnan and TimeStepCalculatorLocal objects
and its flux templates through a restatement of SolverExplicit::timeStep_
(SolverExplicit.cpp:259-378).  On one thread the oracle's RK4 iterations must
end in the same state bit for bit; on several threads the atomic adds reorder
the face sums of both, so the states agree to rounding.  The binary is built
from the reference tree (make -C oracle ref ref-time); where it is absent the
test is skipped.
"""
import os
import subprocess

import numpy as np
import pytest

from bes_amd import core, meshgen as mg
from oracle import oracle as orc

REF_TIME = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                        "bes_ref_time")

pytestmark = pytest.mark.skipif(not os.path.exists(REF_TIME), reason="reference timing driver not built")


@pytest.mark.parametrize("threads", [1, 4])
def test_oracle_explicit_loop_matches_reference_units(tmp_path, threads):
    md = mg.swept_wing(8, 4, 6)
    mach = core.wing_freestream(0.84, 3.06)
    msh, besi, state = tmp_path / "w.msh", tmp_path / "w.besi", tmp_path / "w.state"
    mg.write_msh2(str(msh), md)
    iters = 5
    mg.write_besi(str(besi), str(msh), 3, mach=mach, cfl=0.8, nsteps=iters, output=str(tmp_path))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run([REF_TIME, str(besi), str(iters), str(state)], env=env, check=True, capture_output=True,
                   cwd=str(tmp_path))
    Uref = np.fromfile(str(state)).reshape(md.ncell, 5)
    finf = core.freestream_primitive(mach, 3)
    U0 = np.tile(core.conserved_from_primitive(finf, 3), (md.ncell, 1))
    U, _ = orc.Mesh(md).explicit(U0, iters, 0.8, finf, nthreads=threads)
    if threads == 1:
        assert np.array_equal(U, Uref)
    else:
        assert np.abs(U - Uref).max() <= 1e-14 * np.abs(Uref).max()
    assert not np.array_equal(U, U0)
