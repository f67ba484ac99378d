"""Converged lift and drag of every solver form against the oracle's FP64 solve.

A 9,216-tet swept wing (M 0.84, alpha 3.06 deg) is converged to a 1e-10 drop
of the density residual by the oracle (the reference's implicit loop on the
CPU in FP64: SolverImplicit.cpp:552-600, with a multicolour ILU(0)-GMRES solve
standing in for AMGCL) and on the GPU by the plain FP64 form and by each
accelerated form (FP32 inner solves with ILU(0), the aggregation AMG and the
learned operator; BiCGStab).  The steady state does not depend on the path,
so every form's force coefficients (the wing's pressure force, the quantity
the reference writes per face to slip_wall_*.csv, SolutionSnapshot.cpp:268-308)
must agree with the oracle's to 1e-6.
"""
import numpy as np
import pytest

from bes_amd import core, meshgen as mg
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MACH, ALPHA = 0.84, 3.06
SCHEDULE = (1.0, 1000.0, (0.5, 1.5))
DROP = 1e-10
TOL = 1e-6


@pytest.fixture(scope="module")
def wing():
    md = mg.swept_wing(16, 8, 12)
    finf = core.freestream_primitive(core.wing_freestream(MACH, ALPHA), 3)
    U0 = np.tile(core.conserved_from_primitive(finf, 3), (md.ncell, 1))
    om = orc.Mesh(md)
    fc = om.faces()
    s = orc.Implicit(om, rtol=5e-2, maxiter=300)
    s.start(*SCHEDULE, stop_drop=DROP)
    U = np.ascontiguousarray(U0.copy())
    hist = s.steps(U, 400, finf)
    drop = hist[-1, 5] / hist[0, 5]
    cl, cd = mg.force_coefficients(fc["S"], fc["center"], fc["btype"], fc["owner"], U, MACH, ALPHA,
                                   md.wing["spec"])[:2]
    print("oracle", md.ncell, "cells:", len(hist), "steps, drop", drop, "CL", cl, "CD", cd, flush=True)
    assert drop <= DROP
    return dict(md=md, finf=finf, U0=U0, fc=fc, CL=cl, CD=cd)


FORMS = {
    "plain_fp64_ilu0": (64, "ilu0", "gmres"),
    "mixed_fp32_ilu0": (32, "ilu0", "gmres"),
    "mixed_fp32_amg": (32, "amg", "gmres"),
    "mixed_fp32_learned": (32, "learned", "gmres"),
    "plain_fp64_bicgstab": (64, "ilu0", "bicgstab"),
}


@pytest.mark.parametrize("form", sorted(FORMS))
def test_converged_forces_match_oracle(wing, form):
    precision, precond, solver = FORMS[form]
    print("case", form, flush=True)
    md, fc = wing["md"], wing["fc"]
    with core.Context(md) as c:
        c.set_freestream(wing["finf"])
        c.set_state(wing["U0"])
        c.implicit_setup(precision, precond, rtol=5e-2, maxiter=300)
        c.implicit_set_solver(solver)
        if precond == "learned":
            c.lp_init(20261202)
        if precond == "amg":
            c.amg_set_params(relax="block_jacobi")
        hist = c.implicit_run(400, *SCHEDULE, stop_drop=DROP, repeat_nv=False)
        U = c.get_state()
    drop = hist[-1, 5] / hist[0, 5]
    cl, cd = mg.force_coefficients(fc["S"], fc["center"], fc["btype"], fc["owner"], U, MACH, ALPHA,
                                   md.wing["spec"])[:2]
    print(form, len(hist), "steps, drop", drop, "dCL", cl - wing["CL"], "dCD", cd - wing["CD"])
    assert drop <= DROP
    assert abs(cl - wing["CL"]) <= TOL and abs(cd - wing["CD"]) <= TOL
