"""The CPU oracle against golden vectors produced by the reference itself.

tests/golden/*.npz were written by tests/golden/make_golden.py, which runs the
reference's own compiled translation units (Mesh, Element, GradientGreenGauss,
LimiterVenkatakrishnan, TimeStepCalculator*, Flux* templates, Eigen AD).  The
oracle restates those algorithms in the reference's evaluation order, so the
comparison is bitwise.
"""
import numpy as np

from conftest import golden_mesh
from oracle import oracle as orc


def _setup(g):
    md = golden_mesh(g)
    return md, orc.Mesh(md, orc.COMPAT), int(g["ndim"])


def test_geometry_bitwise(golden):
    _, g = golden
    _, m, nd = _setup(g)
    F, C = m.faces(), m.cells()
    assert np.array_equal(F["owner"], g["face_owner"])
    assert np.array_equal(F["nbr"], g["face_nbr"])
    assert np.array_equal(F["btype"], g["face_btype"])
    for a, b in [("S", "face_S"), ("n", "face_n"), ("area", "face_area"), ("center", "face_center")]:
        assert np.array_equal(F[a], g[b]), a
    for a, b in [("vol", "cell_vol"), ("centroid", "cell_centroid"), ("w", "cell_w"), ("r", "cell_r"),
                 ("cface", "cell_faces"), ("cadj", "cell_adj"), ("rsel", "cell_rsel")]:
        assert np.array_equal(C[a], g[b]), a


def test_reconstruction_lookup_quirk(golden):
    """Element.cpp:102-105 keys the face map before ids exist: only global face 0 hits."""
    _, g = golden
    k = g["cell_faces"].shape[1]
    pred = np.where(g["cell_faces"] == 0, k - 1, 0)
    assert np.array_equal(g["cell_rsel"], pred)


def test_fields_bitwise(golden):
    _, g = golden
    _, m, nd = _setup(g)
    U = g["state"]
    assert np.array_equal(m.gradient(U), g["grad"])
    assert np.array_equal(m.limiter(U, g["grad"]), g["lim"])
    assert np.array_equal(m.timestep(U, float(g["cfl"]), True), g["dt_local"])
    assert np.array_equal(m.timestep(U, float(g["cfl"]), False)[:1], g["dt_global"])


def test_residual_bitwise(golden):
    _, g = golden
    _, m, nd = _setup(g)
    finf = orc.freestream_primitive(g["mach"], nd)
    R, F = m.residual(g["state"], g["grad"], g["lim"], finf, with_flux=True)
    assert np.array_equal(F, g["face_flux"])
    assert np.array_equal(R, g["resid"])


def test_explicit_iterations_bitwise(golden):
    _, g = golden
    _, m, nd = _setup(g)
    finf = orc.freestream_primitive(g["mach"], nd)
    U, linf = m.explicit(g["state"], int(g["niter"]), float(g["cfl"]), finf, nthreads=1)
    assert np.array_equal(U, g["explicit_state"])
    assert np.array_equal(linf, g["explicit_linf"])


def test_implicit_assembly_bitwise(golden):
    _, g = golden
    _, m, nd = _setup(g)
    nv = nd + 2
    finf = orc.freestream_primitive(g["mach"], nd)
    rp, col, val, rhs, _ = m.implicit_assemble(g["state"], float(g["cfl"]), finf)
    assert np.array_equal(rhs, g["impl_rhs"])
    # reference scalar CSR: row (e, a) holds blocks of row e in order, row a of each
    csr = np.concatenate([val[rp[e]:rp[e + 1], a, :].ravel() for e in range(m.ncell) for a in range(nv)])
    cols = np.concatenate([(col[rp[e]:rp[e + 1], None] * nv + np.arange(nv)).ravel()
                           for e in range(m.ncell) for a in range(nv)])
    assert np.array_equal(csr, g["csr_val"])
    assert np.array_equal(cols, g["csr_col"])


def test_flux_and_jacobian_kats_bitwise(golden):
    _, g = golden
    nd = int(g["ndim"])
    nv = nd + 2
    kin = g["kat_in"]
    for kind in (1, 2, 3):
        sel = kin[:, 0] == kind
        r = kin[sel]
        UL, UR, n = r[:, 1:1 + nv], r[:, 1 + nv:2 + 2 * nv], r[:, 2 + 2 * nv:]
        assert np.array_equal(orc.flux(nd, kind, UL, UR, n), g["kat_flux"][sel])
        JL, JR = orc.jacobian(nd, kind, UL, UR, n)
        assert np.array_equal(JL, g["kat_JL"][sel])
        assert np.array_equal(JR, g["kat_JR"][sel])


def test_kat_coverage(golden):
    """The known-answer inputs reach every far-field branch and the Harten fix."""
    _, g = golden
    nd = int(g["ndim"])
    nv = nd + 2
    kin = g["kat_in"]
    ff = kin[kin[:, 0] == 3]
    UL, n = ff[:, 1:1 + nv], ff[:, 2 + 2 * nv:]
    rho = UL[:, 0]
    u = UL[:, 1:1 + nd] / rho[:, None]
    p = 0.4 * (UL[:, -1] - 0.5 * rho * (u * u).sum(1))
    c = np.sqrt(1.4 * p / rho)
    vn = (u * n).sum(1)
    branches = {(bool(a), bool(b)) for a, b in zip(np.abs(vn) / c >= 1, vn <= 0)}
    assert len(branches) == 4
    roe = kin[kin[:, 0] == 1]
    ul, ur, nn = roe[:, 1:1 + nv], roe[:, 1 + nv:1 + 2 * nv], roe[:, 2 + 2 * nv:]
    srl, srr = np.sqrt(ul[:, 0]), np.sqrt(ur[:, 0])
    ua = (ul[:, 1:1 + nd] / srl[:, None] + ur[:, 1:1 + nd] / srr[:, None]) / (srl + srr)[:, None]
    vna = np.abs((ua * nn).sum(1))
    assert (vna < 0.05).any()
