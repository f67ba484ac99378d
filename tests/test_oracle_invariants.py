"""Checks the mathematics fixes, independent of the reference's code.

These pin the oracle where golden vectors cannot: closed forms (physical flux,
slip-wall pressure flux), consistency and symmetry of the Roe flux,
finite-difference Jacobians, freestream preservation and discrete
conservation of the residual, exact geometry of tetrahedra.
"""
import numpy as np
import pytest

from bes_amd import meshgen as mg
from oracle import oracle as orc

G, RG = orc.GAMMA, orc.GAS_R


def phys_flux(U, n):
    nd = len(n)
    rho = U[0]
    u = U[1:1 + nd] / rho
    p = (G - 1) * (U[-1] - 0.5 * rho * u.dot(u))
    V = u.dot(n)
    return np.concatenate([[rho * V], rho * u * V + p * n, [(U[-1] + p) * V]])


def rand_state(rng, nd, mach=0.6):
    P = np.zeros(nd + 3)
    P[0] = rng.uniform(0.7, 1.3)
    P[1:1 + nd] = rng.uniform(-mach, mach, nd)
    P[nd + 1] = rng.uniform(0.5, 0.9)
    P[nd + 2] = P[nd + 1] / (P[0] * RG)
    return orc.conserved_from_primitive(P, nd)


def unit(rng, nd):
    v = rng.normal(size=nd)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("nd", [2, 3])
def test_roe_consistency_and_symmetry(nd):
    rng = np.random.default_rng(1)
    for _ in range(20):
        U, W, n = rand_state(rng, nd), rand_state(rng, nd), unit(rng, nd)
        F = orc.flux(nd, 1, U, U, n)[0]
        np.testing.assert_allclose(F, phys_flux(U, n), rtol=1e-13, atol=1e-13)
        F1 = orc.flux(nd, 1, U, W, n)[0]
        F2 = orc.flux(nd, 1, W, U, -n)[0]
        np.testing.assert_allclose(F1, -F2, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("nd", [2, 3])
def test_slip_and_farfield_closed_forms(nd):
    rng = np.random.default_rng(2)
    for _ in range(10):
        U, n = rand_state(rng, nd), unit(rng, nd)
        rho = U[0]
        u = U[1:1 + nd] / rho
        p = (G - 1) * (U[-1] - 0.5 * rho * u.dot(u))
        F = orc.flux(nd, 2, U, np.zeros(nd + 3), n)[0]
        np.testing.assert_allclose(F, np.concatenate([[0.0], p * n, [0.0]]), rtol=1e-14, atol=1e-15)
    for mach in (0.3, 0.84, 1.7):
        for _ in range(8):
            m = mach * unit(rng, nd)
            Pinf = orc.freestream_primitive(m, nd)
            Uinf = orc.conserved_from_primitive(Pinf, nd)
            n = unit(rng, nd)
            F = orc.flux(nd, 3, Uinf, Pinf, n)[0]
            np.testing.assert_allclose(F, phys_flux(Uinf, n), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("nd,kind", [(2, 1), (3, 1), (2, 2), (3, 2), (2, 3), (3, 3)])
def test_jacobian_matches_finite_differences(nd, kind):
    rng = np.random.default_rng(3 + nd + kind)
    nv = nd + 2
    for _ in range(10):
        UL, n = rand_state(rng, nd), unit(rng, nd)
        if kind == 3:
            UR = orc.freestream_primitive(rng.choice([0.4, 1.8]) * unit(rng, nd), nd)
        else:
            UR = np.concatenate([rand_state(rng, nd), [0.0]])
        JL, JR = orc.jacobian(nd, kind, UL, UR, n)
        for side, J in ((0, JL[0]), (1, JR[0])):
            if side == 1 and kind != 1:
                continue
            fd = np.zeros((nv, nv))
            for i in range(nv):
                h = 1e-6 * max(1.0, abs((UL if side == 0 else UR)[i]))
                a, b = UL.copy(), UL.copy()
                ra, rb = UR.copy(), UR.copy()
                if side == 0:
                    a[i] += h; b[i] -= h
                else:
                    ra[i] += h; rb[i] -= h
                fd[:, i] = (orc.flux(nd, kind, a, ra, n)[0] - orc.flux(nd, kind, b, rb, n)[0]) / (2 * h)
            # away from the Harten/branch kinks central differences agree to ~1e-7
            np.testing.assert_allclose(J, fd, rtol=1e-5, atol=1e-6)


def test_freestream_preservation_and_conservation():
    md = mg.box_tets(3, 3, 3, seed=4, slip_faces=())
    m = orc.Mesh(md)
    mach = (0.5, 0.2, -0.1)
    Pinf = orc.freestream_primitive(mach, 3)
    U = np.tile(orc.conserved_from_primitive(Pinf, 3), (m.ncell, 1))
    Gr = m.gradient(U)
    phi = m.limiter(U, Gr)
    R = m.residual(U, Gr, phi, Pinf)
    assert np.abs(R).max() < 1e-13
    # discrete conservation on a perturbed state: interior fluxes cancel
    U2 = orc.perturbed_state(m.ncell, 3, mach, seed=9)
    Gr = m.gradient(U2)
    phi = m.limiter(U2, Gr)
    R, F = m.residual(U2, Gr, phi, Pinf, with_flux=True)
    fa = m.faces()
    bnd = fa["nbr"] < 0
    total = (F[bnd] * fa["area"][bnd, None]).sum(0)
    np.testing.assert_allclose(R.sum(0), total, rtol=1e-12, atol=1e-13)
    assert (phi >= 0).all() and (phi <= 1).all()


def test_constant_field_zero_gradient_interior():
    md = mg.box_tets(3, 3, 3, seed=5)
    m = orc.Mesh(md)
    U = np.tile([1.1, 0.3, -0.2, 0.1, 2.0], (m.ncell, 1))
    Gr = m.gradient(U)
    c = m.cells()
    fa = m.faces()
    noslip = ~(fa["btype"][c["cface"]] == 2).any(1)
    scale = np.abs(U).max() * np.abs(fa["S"]).max() / c["vol"].min()
    assert np.abs(Gr[noslip]).max() < 1e-13 * scale


def test_exact_geometry_mode():
    md = mg.box_tets(3, 2, 2, seed=6, extent=(1.0, 2.0, 0.5))
    me = orc.Mesh(md, orc.EXACT).cells()
    mc = orc.Mesh(md, orc.COMPAT).cells()
    vol6 = np.abs(mg.tet_signed_volume6(md.nodes, md.cells))
    np.testing.assert_allclose(me["vol"], vol6 / 6, rtol=1e-13)
    np.testing.assert_allclose(mc["vol"], vol6 / 2, rtol=1e-13)   # compat: 3x the volume
    np.testing.assert_allclose(me["vol"].sum(), 1.0, rtol=1e-13)
    np.testing.assert_allclose(me["centroid"], md.nodes[md.cells].mean(1), rtol=1e-13, atol=1e-15)
    k = me["rsel"].shape[1]
    assert np.array_equal(me["rsel"], np.tile(np.arange(k), (md.ncell, 1)))
