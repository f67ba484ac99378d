"""Solution files in the reference's formats, written from the device state.

Legacy VTK and CSV of the cell primitives and the slip-wall CSV follow
SolutionSnapshot.cpp:146-308 with the reference's conversion to its
"dimensional" variables [rho, u/c, gamma p, gamma R T] (GasPerfect.tpp:294-320).
The conversion is pinned by its closed form at the freestream, where the
non-dimensionalisation of Config.cpp:479-488 makes the file hold rho = 1, the
freestream Mach vector, gamma p = 1 and gamma R T = 1; connectivity, node
coordinates, centroids and wall area vectors are compared with the mesh and
the geometry the library reports.  The reference prints with the stream's
default six significant digits, so values agree to that.
"""
import numpy as np
import pytest

from bes_amd import core, meshgen as mg
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MACH3 = (0.84 * np.cos(np.radians(3.06)), 0.0, 0.84 * np.sin(np.radians(3.06)))
RT6 = dict(rtol=6e-6, atol=1e-300)   # six significant digits


def read_vtk(path):
    """Sections of a legacy ASCII unstructured-grid file, as SolutionSnapshot.cpp:172-266 writes it."""
    lines = [l.split() for l in open(path)]
    assert lines[0] == ["#", "vtk", "DataFile", "Version", "2.0"] and lines[2] == ["ASCII"]
    assert lines[3] == ["DATASET", "UNSTRUCTURED_GRID"]
    out, i = {}, 4
    while i < len(lines):
        h = lines[i]
        if h[0] == "POINTS":
            n = int(h[1])
            out["points"] = np.array(lines[i + 1:i + 1 + n], dtype=np.float64)
            i += 1 + n
        elif h[0] == "CELLS":
            n = int(h[1])
            rows = lines[i + 1:i + 1 + n]
            assert sum(len(r) for r in rows) == int(h[2])
            out["cells"] = np.array(rows, dtype=np.int64)
            i += 1 + n
        elif h[0] == "CELL_TYPES":
            n = int(h[1])
            out["types"] = np.array([r[0] for r in lines[i + 1:i + 1 + n]], dtype=np.int64)
            i += 1 + n
        elif h[0] == "CELL_DATA":
            out["ncell"] = int(h[1])
            i += 1
        elif h[0] == "SCALARS":
            assert lines[i + 1] == ["LOOKUP_TABLE", "default"]
            n = out["ncell"]
            out[h[1]] = np.array([r[0] for r in lines[i + 2:i + 2 + n]], dtype=np.float64)
            i += 2 + n
        elif h[0] == "VECTORS":
            n = out["ncell"]
            out[h[1]] = np.array(lines[i + 1:i + 1 + n], dtype=np.float64)
            i += 1 + n
        else:
            raise AssertionError(f"unexpected line {i}: {h}")
    return out


def read_csv(path):
    rows = [l.rstrip("\n") for l in open(path) if l.strip()]
    assert all(r.endswith(",") for r in rows)   # every field is followed by a comma
    return np.array([[float(x) for x in r[:-1].split(",")] for r in rows])


def dimensional(U, nd, gamma=orc.GAMMA, R=orc.GAS_R):
    """[rho, u/c, gamma p, gamma R T] of conserved states (GasPerfect.tpp:34-111, 294-320)."""
    rho, m, E = U[:, 0], U[:, 1:1 + nd], U[:, nd + 1]
    e = E - (m * m).sum(1) / (2.0 * rho)
    p = (gamma - 1.0) * e
    T = e / (R / (gamma - 1.0) * rho)
    c = np.sqrt(gamma * R * T)
    return np.column_stack([rho, m / rho[:, None] / c[:, None], gamma * p, gamma * R * T])


def _check_vtk(v, md, P):
    nd = md.ndim
    pts = v["points"]
    assert pts.shape == (md.nnode, 3)
    np.testing.assert_allclose(pts[:, :nd], md.nodes, rtol=6e-6, atol=1e-6 * np.abs(md.nodes).max())
    if nd == 2:
        assert np.all(pts[:, 2] == 0)
    assert np.array_equal(v["cells"][:, 0], np.full(md.ncell, nd + 1))
    assert np.array_equal(v["cells"][:, 1:], md.cells)
    assert np.all(v["types"] == (5 if nd == 2 else 10))
    assert v["ncell"] == md.ncell
    np.testing.assert_allclose(v["rho_nondim"], P[:, 0], **RT6)
    np.testing.assert_allclose(v["P_nondim"], P[:, nd + 1], **RT6)
    np.testing.assert_allclose(v["T_nondim"], P[:, nd + 2], **RT6)
    assert v["M"].shape == (md.ncell, 3)
    np.testing.assert_allclose(v["M"][:, :nd], P[:, 1:1 + nd], rtol=6e-6, atol=1e-6 * np.abs(P[:, 1:1 + nd]).max())
    if nd == 2:
        assert np.all(v["M"][:, 2] == 0)


def test_freestream_closed_form_3d(tmp_path):
    md = mg.swept_wing(8, 4, 6)
    finf = orc.freestream_primitive(MACH3, 3)
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(np.tile(orc.conserved_from_primitive(finf, 3), (md.ncell, 1)))
        c.write_solution(tmp_path / "s.vtk", "vtk")
        c.write_solution(tmp_path / "s.csv", "csv")
        c.write_slip_wall(tmp_path / "w.csv")
        geo = c.geometry()
    closed = np.tile(np.r_[1.0, MACH3, 1.0, 1.0], (md.ncell, 1))
    _check_vtk(read_vtk(tmp_path / "s.vtk"), md, closed)
    csv = read_csv(tmp_path / "s.csv")
    assert csv.shape == (md.ncell, 3 + 6)
    np.testing.assert_allclose(csv[:, :3], geo["centroid"], rtol=6e-6, atol=1e-6)
    np.testing.assert_allclose(csv[:, 3:], closed, rtol=6e-6, atol=1e-7)
    wall = read_csv(tmp_path / "w.csv")
    faces = np.flatnonzero(geo["btype"] == 2)   # slip-wall faces in the reference's face order
    assert len(faces) > 0 and wall.shape == (len(faces), 3 + 3 + 6)
    S = geo["S"][faces]
    np.testing.assert_allclose(wall[:, 3:6], S, rtol=6e-6, atol=1e-6 * np.abs(S).max())
    np.testing.assert_allclose(wall[:, 6:], closed[:len(faces)], rtol=6e-6, atol=1e-7)


@pytest.mark.parametrize("nd", [2, 3])
def test_perturbed_state_files(tmp_path, nd):
    md = mg.swept_wing(8, 4, 6) if nd == 3 else mg.bump_channel(20, 10, seed=4)
    mach = MACH3 if nd == 3 else (0.6, 0.03)
    finf = orc.freestream_primitive(mach, nd)
    U = orc.perturbed_state(md.ncell, nd, mach, seed=11, amp=0.05)
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U)
        c.write_solution(tmp_path / "s.vtk", "vtk")
        c.write_solution(tmp_path / "s.csv", "csv")
        c.write_slip_wall(tmp_path / "w.csv")
        geo = c.geometry()
    P = dimensional(U, nd)
    _check_vtk(read_vtk(tmp_path / "s.vtk"), md, P)
    csv = read_csv(tmp_path / "s.csv")
    np.testing.assert_allclose(csv[:, nd:], P, rtol=6e-6, atol=1e-6 * np.abs(P).max())
    wall = read_csv(tmp_path / "w.csv")
    faces = np.flatnonzero(geo["btype"] == 2)
    own = geo["owner"][faces]
    m = np.argmax(geo["cface"][own] == faces[:, None], axis=1)    # owner's local index of the face
    centre = geo["centroid"][own] + geo["r"][own, m]
    np.testing.assert_allclose(wall[:, :nd], centre, rtol=6e-6, atol=1e-6 * np.abs(centre).max())
    np.testing.assert_allclose(wall[:, nd:2 * nd], geo["S"][faces], rtol=6e-6,
                               atol=1e-6 * np.abs(geo["S"][faces]).max())
    np.testing.assert_allclose(wall[:, 2 * nd:], P[own], rtol=6e-6,
                               atol=1e-6 * np.abs(P).max())


def test_batched_condition_selects_state(tmp_path):
    md = mg.swept_wing(8, 4, 6)
    machs = [MACH3, (0.5, 0.0, 0.1)]
    fb = np.stack([orc.freestream_primitive(m, 3) for m in machs])
    with core.Context(md) as c:
        c.set_batch(2)
        c.set_freestream(fb)
        c.set_state(np.stack([np.tile(orc.conserved_from_primitive(f, 3), (md.ncell, 1)) for f in fb]))
        for b in range(2):
            c.write_solution(tmp_path / f"s{b}.csv", "csv", cond=b)
        with pytest.raises(core.BesError):
            c.write_solution(tmp_path / "bad.csv", "csv", cond=2)
        with pytest.raises(core.BesError):
            c.write_slip_wall(tmp_path / "nodir" / "w.csv")
    for b in range(2):
        csv = read_csv(tmp_path / f"s{b}.csv")
        np.testing.assert_allclose(csv[:, 3:], np.tile(np.r_[1.0, machs[b], 1.0, 1.0], (md.ncell, 1)),
                                   rtol=6e-6, atol=1e-7)
