"""Pins of the learned-preconditioner restatement (oracle/gnn.py) on the CPU.

The network has no counterpart in the reference, so its restatement is pinned
by what the mathematics fixes: central finite differences of the float64 loss
for the back-propagated gradient, the two ends of the operator family (block
Jacobi and the first-order Neumann operator), the exact transpose of the block
products, and the closed form of one Richardson step.
"""
import numpy as np
import pytest

from bes_amd import meshgen as mg
from oracle import gnn, oracle as orc


@pytest.fixture(scope="module")
def system():
    md = mg.swept_wing(8, 4, 6)
    mvec = (0.84 * np.cos(np.radians(3.06)), 0.0, 0.84 * np.sin(np.radians(3.06)))
    finf = orc.freestream_primitive(mvec, 3)
    U = orc.perturbed_state(md.ncell, 3, mvec, seed=3, amp=0.02)
    om = orc.Mesh(md)
    g = gnn.Graph.from_geometry({**om.faces(), **om.cells()})
    rp, col, val, rhs, _ = om.implicit_assemble(U, 10.0, finf)
    D, O = gnn.blocks_from_bsr(rp, col, val, md.ncell, 4, g.nbr)
    return g, D, O, rhs, (rp, col, val)


def test_graph_is_symmetric(system):
    g = system[0]
    for i in range(g.N):
        for m in range(g.k):
            j = g.nbr[i, m]
            if j >= 0:
                assert g.nbr[j, g.mrev[i, m]] == i


def test_block_products_match_csr(system):
    g, D, O, rhs, (rp, col, val) = system
    import scipy.sparse as sp
    A = sp.bsr_matrix((val, col, rp), shape=(g.N * 5, g.N * 5)).tocsr()
    rng = np.random.default_rng(1)
    v = rng.standard_normal((g.N, 5))
    w = rng.standard_normal((g.N, 5))
    assert np.allclose(gnn.block_apply(g, D, O, v).ravel(), A @ v.ravel(), rtol=1e-12, atol=1e-12)
    assert np.allclose(gnn.block_apply_t(g, D, O, w).ravel(), A.T @ w.ravel(), rtol=1e-12, atol=1e-12)


def test_operator_family_ends(system):
    g, D, O, rhs, _ = system
    rng = np.random.default_rng(2)
    v = rng.standard_normal((g.N, 5))
    Dinv = np.linalg.inv(D)
    bj = np.einsum("nij,nj->ni", Dinv, v)
    # G -> 0: block Jacobi; G -> 1: D^-1 v - D^-1 O D^-1 v (two block-Jacobi sweeps)
    for big, expect in ((-60.0, bj), (60.0, bj - np.einsum("nij,nj->ni", Dinv,
                                                           gnn.block_apply(g, np.zeros_like(D), O, bj)))):
        Md, Me = gnn.operator_blocks(g, D, O, np.full((g.N, g.k, 5), big))
        assert np.allclose(gnn.block_apply(g, Md, Me, v), expect, rtol=1e-12, atol=1e-12)


def test_one_richardson_step_closed_form(system):
    g, D, O, rhs, _ = system
    rng = np.random.default_rng(3)
    logit = rng.standard_normal((g.N, g.k, 5))
    Md, Me = gnn.operator_blocks(g, D, O, logit)
    V = rng.standard_normal((2, g.N, 5))
    L, _ = gnn.richardson_loss(g, D, O, Md, Me, V, 1)
    ref = np.mean([np.sum((v - gnn.block_apply(g, D, O, gnn.block_apply(g, Md, Me, v))) ** 2) / np.sum(v * v)
                   for v in V])
    assert abs(L - ref) <= 1e-12 * ref


@pytest.mark.parametrize("kind,mr", [("richardson", False), ("richardson", True), ("gmres", False)])
def test_gradient_matches_finite_differences(system, kind, mr):
    g, D, O, rhs, _ = system
    rng = np.random.default_rng(0)
    theta = gnn.init_params(7).astype(np.float64) + 0.05 * rng.standard_normal(gnn.param_count())
    V = rng.standard_normal((2, g.N, 5))
    kw = dict(K=3, bf16_mode=False, mr=mr, kind=kind)
    L, gr = gnn.loss_and_grad(theta, g, D, O, rhs, V, **kw)
    assert abs(L - gnn.loss(theta, g, D, O, rhs, V, **kw)) <= 1e-12 * L
    # parameters of every group: encoder, each layer, decoder
    layout = gnn.param_layout()
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in layout])
    picks = []
    for q, (name, shape) in enumerate(layout):
        lo, hi = offs[q], offs[q + 1]
        cand = [p for p in rng.permutation(np.arange(lo, hi)) if gr[p] != 0.0][:2]
        picks += cand
    assert len(picks) >= 12
    for p in picks:
        e = np.zeros_like(theta)
        e[p] = 1e-6
        fd = (gnn.loss(theta + e, g, D, O, rhs, V, **kw) - gnn.loss(theta - e, g, D, O, rhs, V, **kw)) / 2e-6
        assert abs(fd - gr[p]) <= 1e-5 * max(abs(fd), 1e-3), (p, fd, gr[p])


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0, 0.0])
    r = gnn.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125   # ties to even, then up
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)
    assert r[-1] == 0.0


def test_init_params_layout():
    p = gnn.init_params(11)
    assert p.dtype == np.float32 and p.size == gnn.param_count() == 15616
    u = gnn.unpack(p)
    assert np.all(u["We"] == 0) and np.all(u["ce"][:5] == gnn.CE0) and np.all(u["ce"][5:] == 0)
    assert np.all(u["W0"][:, 31:] == 0)
    lim = np.sqrt(6.0 / 96)
    assert np.abs(u["W1_0"]).max() <= lim and np.abs(u["W2_0"]).max() <= 0.5 * lim
    assert np.array_equal(p, gnn.init_params(11)) and not np.array_equal(p, gnn.init_params(12))
