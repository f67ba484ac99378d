"""CPU restatement of the learned block preconditioner (test infrastructure).

Used only by tests/ and as the checker of the HIP kernels in
bes_amd/csrc/learned.hip; it is never the computation that ships.  The
reference has no learned component: this network is an addition whose job is
to propose a better Krylov preconditioner for the implicit system that the
reference assembles (SolverImplicit.cpp:358-507) and hands to AMGCL
(CoefficientMatrix.cpp:210-325).

The operator.  For the block system A x = b (nv x nv blocks: the diagonal
block D_i and one off-diagonal block O_im per internal face m of cell i,
j = nbr(i,m)), the preconditioner is

    (M v)_i = D_i^-1 v_i - sum_m G_im P_im v_j,    P_im = D_i^-1 O_im D_j^-1,
    G_im = diag(sigmoid(s_im)),  s_im in R^nv from the network.

Every G between 0 (block Jacobi) and I (the first-order Neumann approximation
of A^-1 about the block diagonal, i.e. two block-Jacobi sweeps) is in the
family; the network picks a per-face, per-row blend.  M has the sparsity of A
and is linear in v, so it is a fixed right preconditioner of GMRES; zero
decoder weights give G = I/2 everywhere.

The network (message passing on the cell dual graph, width 32):
    x_i   = [D_i/dmax_i (nv*nv), b_i/bmax (nv), 1, 0...]           (32)
    a_im  = [O_im/dmax_i (nv*nv), n_im (nd), area_im/V_i^(2/3), 0...] (32)
    h0    = relu(W0 x + c0)
    layer l = 1..3:
        msg_im = relu(W1 [h_nbr ; a_im] + c1)                        (64 -> 32)
        agg_i  = sum_m msg_im
        h_i   <- h_i + relu(W2 [h_i ; agg_i] + c2)                   (64 -> 32)
    s_im = We [h_i ; h_nbr] + ce                                      (nv of 32 outputs)
with dmax_i = max |D_i| and bmax the largest |b| per variable.

Loss: K steps of preconditioned Richardson iteration from each probe vector
v_q (the right-hand side and random vectors),
    r_0 = v_q,  r_{k+1} = r_k - A M r_k,   L = (1/Q) sum_q ||r_K||^2 / ||v_q||^2.
Unlike the one-step residual ||v - A M v||, it grows quickly for operators
whose A M has eigenvalues far from 1, the ones that stall GMRES.

Precision model of the GPU kernels (bf16=True): every matrix-product operand
(features, hidden states, aggregated messages, weights and back-propagated
output gradients) is rounded to bfloat16 (round to nearest even); products
accumulate and element-wise work runs in float32 on the GPU (float64 here, so
agreement is to a tolerance).  Hidden states are stored in bfloat16.  With
bf16=False everything is float64, which the finite-difference tests use to pin
the backward formulas.
"""
from __future__ import annotations

import numpy as np

WIDTH = 32
LAYERS = 3
GMAX = 1.0      # blends G = GMAX * sigmoid(s)
CE0 = 3.0       # initial decoder bias: G = sigmoid(3) = 0.953, near the Neumann end


# ---------------------------------------------------------------------------
# parameter layout (flat float32 vector; identical offsets in learned.hip)
# ---------------------------------------------------------------------------

def param_layout():
    """Ordered (name, shape) of the flat parameter vector."""
    W = WIDTH
    out = [("W0", (W, W)), ("c0", (W,))]
    for l in range(LAYERS):
        out += [(f"W1_{l}", (W, 2 * W)), (f"c1_{l}", (W,)), (f"W2_{l}", (W, 2 * W)), (f"c2_{l}", (W,))]
    out += [("We", (W, 2 * W)), ("ce", (W,))]
    return out


def param_count():
    return sum(int(np.prod(s)) for _, s in param_layout())


def unpack(theta):
    p, o = {}, 0
    for name, shape in param_layout():
        n = int(np.prod(shape))
        p[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return p


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in param_layout()])


def _splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


def init_params(seed, nv=5):
    """Deterministic initialisation shared with the library (splitmix64 stream).

    Weights of the encoder and message layers: uniform in +-sqrt(6/(fan_in+fan_out))
    (Glorot); the update layers are scaled by 0.5; biases zero; the decoder
    weights zero and its first nv biases CE0, so the initial blend is
    sigmoid(CE0) on every face.  Columns of the encoder that read padding stay
    zero.
    """
    state = seed & 0xFFFFFFFFFFFFFFFF
    p = {}
    for name, shape in param_layout():
        a = np.zeros(shape)
        if name[0] == "W" and name != "We":
            fan_out, fan_in = shape
            lim = np.sqrt(6.0 / (fan_in + fan_out)) * (0.5 if name.startswith("W2") else 1.0)
            flat = np.empty(a.size)
            for q in range(a.size):
                state, z = _splitmix64(state)
                flat[q] = ((z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * lim
            a = flat.reshape(shape)
        p[name] = a
    nb = nv * nv
    p["W0"][:, nb + nv + 1:] = 0.0
    p["ce"][:nv] = CE0
    return pack(p).astype(np.float32)


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(np.isnan(f), u, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(f.shape)


def _ident(x):
    return np.asarray(x, dtype=np.float64)


# ---------------------------------------------------------------------------
# graph and features
# ---------------------------------------------------------------------------

class Graph:
    """Cell dual graph in reference cell order.

    nbr [N][k]: neighbour through local face m (or < 0 on the boundary);
    mrev [N][k]: local index of the shared face in the neighbour; efeat
    [N][k][nd+1]: outward unit normal and area / V^(2/3) per (cell, face).
    """

    def __init__(self, nbr, mrev, normal, area, vol):
        self.nbr = np.asarray(nbr, dtype=np.int64)
        self.N, self.k = self.nbr.shape
        self.mrev = np.asarray(mrev, dtype=np.int64)
        self.nd = normal.shape[-1]
        self.nv = self.nd + 2
        self.edge = self.nbr >= 0
        ef = np.zeros((self.N, self.k, self.nd + 1))
        ef[..., :self.nd] = normal
        ef[..., self.nd] = area / np.asarray(vol, dtype=np.float64)[:, None] ** (2.0 / 3.0)
        self.efeat = ef

    @staticmethod
    def from_geometry(geo):
        """From core.Context.geometry() / the oracle mesh arrays (reference order)."""
        cface, cadj = geo["cface"], geo["cadj"]
        N, k = cface.shape
        nbr = np.where(cadj >= 0, cadj, -1)
        mrev = np.full((N, k), -1, dtype=np.int64)
        for m in range(k):
            j = nbr[:, m]
            ok = j >= 0
            for mm in range(k):
                hit = ok & (cface[np.where(ok, j, 0), mm] == cface[:, m])
                mrev[hit, m] = mm
        owner = geo["owner"][cface]
        sign = np.where(owner == np.arange(N)[:, None], 1.0, -1.0)
        normal = geo["n"][cface] * sign[..., None]
        area = geo["area"][cface]
        return Graph(nbr, mrev, normal, area, geo["vol"])


def blocks_from_bsr(rp, col, val, N, k, nbr):
    """Split a block-CSR matrix (rows: own block, then neighbours) into D [N][nv][nv]
    and O [N][k][nv][nv] by local face."""
    nv = val.shape[1]
    D = np.zeros((N, nv, nv))
    O = np.zeros((N, k, nv, nv))
    for i in range(N):
        for q in range(rp[i], rp[i + 1]):
            j = col[q]
            if j == i:
                D[i] = val[q]
            else:
                ms = np.nonzero(nbr[i] == j)[0]
                O[i, ms[0]] = val[q]
    return D, O


def features(g, D, O, b, rnd=bf16):
    """Node features x [N][32], edge features a [N][k][32], and dmax [N]."""
    N, k, nv, nd = g.N, g.k, g.nv, g.nd
    nb = nv * nv
    dmax = np.abs(D.reshape(N, nb)).max(1)
    bmax = np.abs(b).max(0)
    bmax = np.where(bmax > 0, bmax, 1.0)
    x = np.zeros((N, WIDTH))
    x[:, :nb] = D.reshape(N, nb) / dmax[:, None]
    x[:, nb:nb + nv] = b / bmax
    x[:, nb + nv] = 1.0
    a = np.zeros((N, k, WIDTH))
    a[..., :nb] = O.reshape(N, k, nb) / dmax[:, None, None]
    a[..., nb:nb + nd + 1] = g.efeat
    a[~g.edge] = 0.0
    return rnd(x), rnd(a), dmax


# ---------------------------------------------------------------------------
# forward / backward
# ---------------------------------------------------------------------------

def _relu(x):
    return np.maximum(x, 0.0)


def forward(theta, g, D, O, b, bf16_mode=True):
    """Network forward; returns (s [N][k][nv] blend logits, cache)."""
    rnd = bf16 if bf16_mode else _ident
    p = unpack(theta)
    W = {n: rnd(v) for n, v in p.items() if n[0] == "W"}
    x, a, dmax = features(g, D, O, b, rnd)
    nbr = np.where(g.edge, g.nbr, 0)
    cache = dict(x=x, a=a, dmax=dmax, W=W, h=[], msg=[], r=[], inp=[], u=[])
    h = rnd(_relu(x @ W["W0"].T + p["c0"]))
    cache["h"].append(h)
    for l in range(LAYERS):
        inp = np.concatenate([h[nbr], a], axis=-1)            # [N][k][64]
        msg = _relu(inp @ W[f"W1_{l}"].T + p[f"c1_{l}"])
        msg = msg * g.edge[..., None]
        agg = msg.sum(1)
        u = np.concatenate([h, rnd(agg)], axis=-1)
        r = _relu(u @ W[f"W2_{l}"].T + p[f"c2_{l}"])
        cache["inp"].append(inp)
        cache["msg"].append(msg)
        cache["u"].append(u)
        cache["r"].append(r)
        h = rnd(h + r)
        cache["h"].append(h)
    pair = np.concatenate([np.repeat(h[:, None, :], g.k, axis=1), h[nbr]], axis=-1)
    logit = (pair @ W["We"].T + p["ce"])[..., :g.nv]
    cache["pair"] = pair
    return logit, cache


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def neumann_blocks(g, D, O):
    """D^-1 [N][nv][nv] and P_im = D_i^-1 O_im D_j^-1 [N][k][nv][nv] (zero off the graph)."""
    Dinv = np.linalg.inv(D)
    nbr = np.where(g.edge, g.nbr, 0)
    P = np.einsum("nij,nmjk,nmkl->nmil", Dinv, O, Dinv[nbr]) * g.edge[..., None, None]
    return Dinv, P


def operator_blocks(g, D, O, logit):
    """Blocks of M: M_ii = D_i^-1, M_im = -diag(sigmoid(s_im)) P_im."""
    Dinv, P = neumann_blocks(g, D, O)
    G = GMAX * sigmoid(logit)
    return Dinv, -G[..., :, None] * P


def block_apply(g, Bd, Be, v):
    """(B v)_i = Bd_i v_i + sum_m Be_im v_nbr(i,m)."""
    nbr = np.where(g.edge, g.nbr, 0)
    y = np.einsum("nij,nj->ni", Bd, v)
    y += np.einsum("nmij,nmj->ni", Be * g.edge[..., None, None], v[nbr])
    return y


def block_apply_t(g, Bd, Be, v):
    """(B^T v)_j = Bd_j^T v_j + sum over faces of j: Be_(i,mrev)^T v_i, i = nbr(j,m)."""
    nbr = np.where(g.edge, g.nbr, 0)
    mrev = np.where(g.edge, g.mrev, 0)
    y = np.einsum("nji,nj->ni", Bd, v)
    Bt = Be[nbr, mrev]                                         # block of row i, column j
    y += np.einsum("nmji,nmj->ni", Bt * g.edge[..., None, None], v[nbr])
    return y


def richardson_loss(g, D, O, Md, Me, V, K, want_grad=False, mr=False):
    """L = mean_q ||r_K||^2/||v_q||^2 with r_{k+1} = r_k - w_k, w_k = A M r_k
    (mr=False), or the minimal-residual step r_{k+1} = r_k - omega_k w_k,
    omega_k = <r_k, w_k>/<w_k, w_k> (mr=True, invariant to the scale of M);
    optionally dL/dMe [N][k][nv][nv] (the diagonal blocks of M are fixed)."""
    Q = len(V)
    nbr = np.where(g.edge, g.nbr, 0)
    L = 0.0
    dMe = np.zeros_like(Me) if want_grad else None
    for v in V:
        vv = float((v * v).sum())
        rs, ws, om = [v], [], []
        for _ in range(K):
            w = block_apply(g, D, O, block_apply(g, Md, Me, rs[-1]))
            o = float((rs[-1] * w).sum()) / float((w * w).sum()) if mr else 1.0
            ws.append(w)
            om.append(o)
            rs.append(rs[-1] - o * w)
        L += float((rs[-1] ** 2).sum()) / vv / Q
        if want_grad:
            gk = 2.0 / (Q * vv) * rs[-1]
            for k in reversed(range(K)):
                w, o, r = ws[k], om[k], rs[k]
                if mr:
                    beta = float((gk * w).sum()) / float((w * w).sum())
                    gw = -beta * (r - 2.0 * o * w) - o * gk
                    gr = gk - beta * w
                else:
                    gw = -gk
                    gr = gk
                at = block_apply_t(g, D, O, gw)                # A^T dL/dw
                dMe += np.einsum("ni,nmj->nmij", at, r[nbr]) * g.edge[..., None, None]
                gk = gr + block_apply_t(g, Md, Me, at)
    return L, dMe


def gmres_loss(g, D, O, Md, Me, V, K, want_grad=False):
    """L = mean_q ||v_q - W y||^2 / ||v_q||^2, W = [A M v_q, (A M)^2 v_q, ...,
    (A M)^K v_q], y the least-squares coefficients: the residual GMRES(K)
    reaches from x0 = 0 on A M x = v_q.  Gradient by the envelope theorem,
    dL/dW = -2 y (x) r, back through the power chain; dL/dMe [N][k][nv][nv]."""
    Q = len(V)
    nbr = np.where(g.edge, g.nbr, 0)
    L = 0.0
    dMe = np.zeros_like(Me) if want_grad else None
    for v in V:
        vv = float((v * v).sum())
        us = [v]
        for _ in range(K):
            us.append(block_apply(g, D, O, block_apply(g, Md, Me, us[-1])))
        W = np.stack([u.ravel() for u in us[1:]])
        y = np.linalg.solve(W @ W.T, W @ v.ravel())
        r = v - (y @ W).reshape(v.shape)
        L += float((r * r).sum()) / vv / Q
        if want_grad:
            gacc = np.zeros_like(v)
            for j in range(K, 0, -1):
                gj = -2.0 / (Q * vv) * y[j - 1] * r + gacc
                at = block_apply_t(g, D, O, gj)
                dMe += np.einsum("ni,nmj->nmij", at, us[j - 1][nbr]) * g.edge[..., None, None]
                gacc = block_apply_t(g, Md, Me, at)
    return L, dMe


def loss_and_grad(theta, g, D, O, b, V, K=3, bf16_mode=True, mr=False, kind="richardson"):
    """Loss over probe vectors V [Q][N][nv] and its gradient (flat, float64)."""
    rnd = bf16 if bf16_mode else _ident
    logit, cache = forward(theta, g, D, O, b, bf16_mode)
    Dinv, P = neumann_blocks(g, D, O)
    sg = sigmoid(logit)
    Md, Me = Dinv, -GMAX * sg[..., :, None] * P
    N, k, nv = g.N, g.k, g.nv
    if kind == "gmres":
        L, dMe = gmres_loss(g, D, O, Md, Me, V, K, True)
    else:
        L, dMe = richardson_loss(g, D, O, Md, Me, V, K, True, mr)
    dlogit = np.zeros((N, k, WIDTH))
    dlogit[..., :nv] = -(dMe * P).sum(-1) * GMAX * sg * (1.0 - sg) * g.edge[..., None]
    dlogit = rnd(dlogit)
    W = cache["W"]
    gr = {name: np.zeros(shape) for name, shape in param_layout()}
    gr["We"] = np.einsum("nmo,nmq->oq", dlogit, cache["pair"])
    gr["ce"] = dlogit.sum((0, 1))
    dpair = dlogit @ W["We"]                                  # [N][k][64]
    dh = dpair[..., :WIDTH].sum(1) + _gather_rev(g, dpair[..., WIDTH:])
    for l in reversed(range(LAYERS)):
        r, u, msg, inp = cache["r"][l], cache["u"][l], cache["msg"][l], cache["inp"][l]
        dpre2 = rnd(dh * (r > 0))
        gr[f"W2_{l}"] = dpre2.T @ u
        gr[f"c2_{l}"] = dpre2.sum(0)
        du = dpre2 @ W[f"W2_{l}"]
        dagg = du[:, WIDTH:]
        dpre1 = rnd(dagg[:, None, :] * (msg > 0) * g.edge[..., None])
        gr[f"W1_{l}"] = np.einsum("nmo,nmq->oq", dpre1, inp)
        gr[f"c1_{l}"] = dpre1.sum((0, 1))
        dinp = dpre1 @ W[f"W1_{l}"]
        dh = dh + du[:, :WIDTH] + _gather_rev(g, dinp[..., :WIDTH])
    h0 = cache["h"][0]
    dpre0 = rnd(dh * (h0 > 0))
    gr["W0"] = dpre0.T @ cache["x"]
    gr["c0"] = dpre0.sum(0)
    return L, pack(gr)


def _gather_rev(g, t):
    """sum over faces m of i of t[nbr(i,m), mrev(i,m)]: what edge (j -> i) sends back to i."""
    nbr = np.where(g.edge, g.nbr, 0)
    mrev = np.where(g.edge, g.mrev, 0)
    return (t[nbr, mrev] * g.edge[..., None]).sum(1)


def loss(theta, g, D, O, b, V, K=3, bf16_mode=True, mr=False, kind="richardson"):
    logit, _ = forward(theta, g, D, O, b, bf16_mode)
    Md, Me = operator_blocks(g, D, O, logit)
    if kind == "gmres":
        return gmres_loss(g, D, O, Md, Me, V, K)[0]
    return richardson_loss(g, D, O, Md, Me, V, K, mr=mr)[0]


class Adam:
    """Adam (Kingma & Ba 2015) on the flat parameters, as the library applies it."""

    def __init__(self, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, theta, grad):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * grad
        self.v = self.b2 * self.v + (1 - self.b2) * grad * grad
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return theta - self.lr * mh / (np.sqrt(vh) + self.eps)
