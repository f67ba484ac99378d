"""Learned preconditioner on the GPU, through the C ABI.

The BF16 matrix-core forward pass, the loss and the back-propagated gradient
are checked against the CPU restatement (oracle/gnn.py, which the CPU tests pin
by finite differences) on the same assembled system, parameters and probe
vectors.  Both sides round the same operands to bfloat16; the GPU accumulates
in float32 and the restatement in float64, so agreement is to a tolerance
written in each test.  The last tests train the network on the GPU (the loss must
fall) and exercise the fallback to ILU(0).
"""
import numpy as np
import pytest

from bes_amd import core, meshgen as mg
from oracle import gnn, oracle as orc

pytestmark = pytest.mark.gpu

MVEC = (0.84 * np.cos(np.radians(3.06)), 0.0, 0.84 * np.sin(np.radians(3.06)))


def blocks_from_scalar_csr(rp, col, val, nv):
    """Block CSR (row pointer, block columns, [nnzb][nv][nv]) from the reference's
    scalar CSR, whose nv rows of one cell share a column pattern: the cell's own
    nv columns, then each neighbour's nv columns (CoefficientMatrix.cpp:114-164)."""
    n = (len(rp) - 1) // nv
    brp = rp[::nv] // (nv * nv)
    bcol = np.empty(brp[-1], dtype=np.int64)
    bval = np.empty((brp[-1], nv, nv))
    for i in range(n):
        nb = brp[i + 1] - brp[i]
        r0 = rp[i * nv]
        bcol[brp[i]:brp[i + 1]] = col[r0:r0 + nb * nv:nv] // nv
        rows = np.stack([val[rp[i * nv + r]:rp[i * nv + r + 1]] for r in range(nv)])   # [nv][nb*nv]
        bval[brp[i]:brp[i + 1]] = rows.reshape(nv, nb, nv).transpose(1, 0, 2)
    return brp, bcol, bval


def _assembled(dims=(8, 4, 6), precision=64, seed=3, cfl=10.0):
    md = mg.swept_wing(*dims)
    finf = orc.freestream_primitive(MVEC, 3)
    U = orc.perturbed_state(md.ncell, 3, MVEC, seed=seed, amp=0.02)
    c = core.Context(md)
    c.set_freestream(finf)
    c.implicit_setup(precision, "learned", restart=30, maxiter=300, rtol=1e-3)
    c.lp_init(20261202)
    c.set_state(U)
    c.implicit_assemble(cfl)
    rp, col, val, rhs = c.implicit_system()
    g = gnn.Graph.from_geometry(c.geometry())
    D, O = gnn.blocks_from_bsr(*blocks_from_scalar_csr(rp, col, val, c.nv), c.ncell, c.k, g.nbr)
    return c, g, D, O, rhs


def _perturbed_params(seed):
    rng = np.random.default_rng(seed)
    p = gnn.init_params(20261202).astype(np.float64)
    p += 0.1 * rng.standard_normal(p.size)
    return p.astype(np.float32)


def test_init_matches_restatement():
    c, g, D, O, rhs = _assembled()
    with c:
        assert np.array_equal(c.lp_params(), gnn.init_params(20261202))


def test_forward_operator_matches_restatement():
    c, g, D, O, rhs = _assembled()
    with c:
        theta = _perturbed_params(1)
        c.lp_set_params(theta)
        Md, Me, G = c.lp_operator()
        logit, _ = gnn.forward(theta.astype(np.float64), g, D, O, rhs, bf16_mode=True)
        Gr = gnn.sigmoid(logit) * g.edge[..., None]
        Mdr, Mer = gnn.operator_blocks(g, D, O, logit)
        # FP64 diagonal inverses
        assert np.allclose(Md, Mdr, rtol=1e-10, atol=1e-12)
        # blends: bf16 products with float32 accumulation against float64
        err = np.abs(G - Gr)[g.edge]
        print("blend error max %.3e mean %.3e" % (err.max(), err.mean()))
        assert err.max() < 2e-2 and err.mean() < 2e-3
        assert np.linalg.norm(Me - Mer) <= 2e-2 * np.linalg.norm(Mer)
        assert np.all(G[~g.edge] == 0) and np.all(Me[~g.edge] == 0)


def test_loss_and_gradient_match_restatement():
    c, g, D, O, rhs = _assembled()
    with c:
        theta = _perturbed_params(2)
        c.lp_set_params(theta)
        rng = np.random.default_rng(5)
        V = np.concatenate([rhs[None], rng.standard_normal((2, c.ncell, c.nv))])
        L, grad = c.lp_loss_grad(V, nrich=3)
        Lr, gr = gnn.loss_and_grad(theta.astype(np.float64), g, D, O, rhs, V, K=3, bf16_mode=True)
        print("loss gpu %.8e cpu %.8e" % (L, Lr))
        assert abs(L - Lr) <= 2e-3 * Lr
        layout = gnn.param_layout()
        offs = np.cumsum([0] + [int(np.prod(s)) for _, s in layout])
        for q, (name, shape) in enumerate(layout):
            a, b = grad[offs[q]:offs[q + 1]].astype(np.float64), gr[offs[q]:offs[q + 1]]
            nb = np.linalg.norm(b)
            rel = np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)
            print(f"{name:6s} |g| {nb:.3e} rel err {rel:.2e}")
            assert rel < 5e-2, name


def _solve_iters(c, precond):
    c.implicit_set_precond(precond)
    it, rr = c.linear_solve(allow_unconverged=True)
    return it, rr


def test_training_lowers_loss_and_reports_iterations():
    """Adam on the GPU lowers the training loss on the assembled system; the
    GMRES iterations of block Jacobi, ILU(0) and the learned operator before
    and after training are reported."""
    md = mg.swept_wing(16, 8, 12)
    finf = orc.freestream_primitive(MVEC, 3)
    U0 = np.tile(orc.conserved_from_primitive(finf, 3), (md.ncell, 1))
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.implicit_setup(64, "jacobi", restart=30, maxiter=600, rtol=1e-3)
        c.lp_init(20261202)
        c.set_state(U0)
        cfl = 10.0
        for step in range(12):
            c.implicit_assemble(cfl)
            c.implicit_set_precond("jacobi")
            c.linear_solve()
            c.implicit_update()
            cfl *= 1.15
        c.implicit_assemble(cfl)
        it_bj, _ = _solve_iters(c, "jacobi")
        it_ilu, _ = _solve_iters(c, "ilu0")
        it_lp0, _ = _solve_iters(c, "learned")
        hist = c.lp_train(40, lr=3e-3, nprobe=4, nrich=3, seed=9)
        it_lp, rr_lp = _solve_iters(c, "learned")
        print(f"system at CFL {cfl:.1f}: block Jacobi {it_bj}, ILU(0) {it_ilu}, learned untrained {it_lp0}, "
              f"after 40 Adam steps {it_lp}; loss {hist[0]:.4f} -> {hist[-1]:.4f}")
        assert np.all(np.isfinite(hist))
        assert np.mean(hist[-5:]) < hist[0]
        assert rr_lp <= 1e-3 * 1.0001
        assert it_lp0 < it_bj


def test_fallback_to_ilu0_when_learned_stalls():
    md = mg.swept_wing(8, 4, 6)
    finf = orc.freestream_primitive(MVEC, 3)
    U = orc.perturbed_state(md.ncell, 3, MVEC, seed=4, amp=0.02)
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.implicit_setup(64, "ilu0", restart=30, maxiter=1000, rtol=1e-8)
        c.set_state(U)
        c.implicit_assemble(50.0)
        it_ilu, _ = c.linear_solve()
        c.lp_init(7)
        c.implicit_set_precond("learned")
        it_lp, _ = c.linear_solve()
        assert it_lp > it_ilu
        # a budget only ILU(0) meets: the learned solve stalls and falls back
        c.implicit_setup(64, "learned", restart=30, maxiter=it_ilu, rtol=1e-8)
        c.lp_init(7)
        c.implicit_assemble(50.0)
        it, rr = c.linear_solve()
        assert rr <= 1e-8 * 1.0001
        assert c.lp_fallbacks() == 1
        assert it == it_ilu + it_ilu   # the stalled budget, then the ILU(0) solve


def test_lagged_operator_rebuild():
    """rebuild interval 3: four assemble-and-solve rounds build the operator
    twice and every solve reaches its tolerance; new parameters force a build."""
    md = mg.swept_wing(8, 4, 6)
    finf = orc.freestream_primitive(MVEC, 3)
    U = orc.perturbed_state(md.ncell, 3, MVEC, seed=6, amp=0.02)
    with core.Context(md) as c:
        c.set_freestream(finf)
        c.set_state(U)
        c.implicit_setup(32, "learned", restart=30, maxiter=300, rtol=1e-3)
        c.lp_init(20261202)
        c.lp_set_rebuild_interval(3)
        for cfl in (10.0, 11.0, 12.0, 13.0):
            c.implicit_assemble(cfl)
            it, rr = c.linear_solve()
            assert rr <= 1e-3 * 1.0001
            c.implicit_update()
        assert c.lp_fallbacks() == 0 and c.lp_builds() == 2
        c.lp_set_params(c.lp_params())
        c.implicit_assemble(13.0)
        c.linear_solve()
        assert c.lp_builds() == 3


def _solve_counts(md, P, U, params, cfls, precision=32):
    """GMRES(30) iterations of the learned form (rebuilt every solve, no
    fallback needed at rtol 5e-2) on the systems of a fixed state at a list of
    Courant numbers."""
    with core.Context(md) as c:
        c.set_freestream(P)
        c.set_state(U)
        c.implicit_setup(precision, "learned", restart=30, maxiter=300, rtol=5e-2)
        c.lp_init(20261202)
        c.lp_set_params(params)
        its = []
        for cfl in cfls:
            c.implicit_assemble(cfl, repeat_nv=False)
            its.append(c.linear_solve(allow_unconverged=True)[0])
        return its, c.lp_fallbacks()


def test_online_training_beats_untrained():
    """The network starts neutral (every blend 1/2, no prior) and is trained on
    the GPU, a few Adam steps on each system the implicit loop assembles on a
    9k-tet wing at M 0.84; on held-out systems (another flow condition, states
    along its own run) the trained operator needs fewer GMRES iterations than
    the untrained one."""
    md = mg.swept_wing(16, 8, 12)
    P = core.freestream_primitive(core.wing_freestream(0.84, 3.06), 3)
    U0 = np.tile(core.conserved_from_primitive(P, 3), (md.ncell, 1))
    with core.Context(md) as c:
        c.set_freestream(P)
        c.set_state(U0)
        c.implicit_setup(32, "learned", restart=30, maxiter=300, rtol=5e-2)
        c.lp_init(20261202)
        c.lp_neutral()
        neutral = c.lp_params()
        c.lp_set_online(2, lr=3e-3, nprobe=4, nrich=3)
        h = c.implicit_run(40, 1.0, 1000.0, (0.5, 1.5), repeat_nv=False)
        trained = c.lp_params()
        print("training run: steps", len(h), "last loss", c.lp_online_loss(), flush=True)
    assert not np.array_equal(trained, neutral)
    # held-out: a different condition; states after 10, 25 and 40 of its own steps (ILU(0) run)
    P2 = core.freestream_primitive(core.wing_freestream(0.75, 1.5), 3)
    states = []
    with core.Context(md) as c:
        c.set_freestream(P2)
        c.set_state(np.tile(core.conserved_from_primitive(P2, 3), (md.ncell, 1)))
        c.implicit_setup(64, "ilu0", rtol=5e-2, maxiter=300)
        for n in (10, 15, 15):
            if not states:
                c.implicit_run(n, 1.0, 1000.0, (0.5, 1.5), repeat_nv=False)
            else:
                c.implicit_continue(n)
            states.append(c.get_state())
    cfls = (50.0, 300.0, 1000.0)
    tot = {"neutral": 0, "trained": 0}
    for U in states:
        for name, p in (("neutral", neutral), ("trained", trained)):
            its, fb = _solve_counts(md, P2, U, p, cfls)
            tot[name] += sum(its)
            print(name, its, "fallbacks", fb, flush=True)
    print(tot)
    assert tot["trained"] < tot["neutral"]
