"""CPU restatement of the AMG preconditioner of the reference's implicit solve
(test infrastructure).

Used only by tests/ as the checker of bes_amd/csrc/amg.hip; never the
computation that ships.  The reference builds an AMGCL solver for every
nonlinear step (CoefficientMatrix.cpp:210-325) from the vendored AMGCL at
commit 801a2fb (include/amgcl__801a2fb), with coarsening "aggregation" and
aggr.block_size = nv (CoefficientMatrix.cpp:231-256), and the relaxation its
.besi names (DAMPED_JACOBI in nacaAirfoil.besi:41-45 and oneraM6.besi:41-45;
SPAI0 is the other diagonal choice, CoefficientMatrix.cpp:258-280).  AMGCL
needs Boost, which is absent, so it is restated here and not run:

  pointwise matrix  one scalar per nv x nv block, its largest |entry|
                    (coarsening/pointwise_aggregates.hpp:188-260)
  strong coupling   eps^2 d_i d_j < a_ij^2 with eps_strong = 0.08 (a float),
                    d the pointwise diagonal (plain_aggregates.hpp:118-135)
  aggregation       rows without a strong coupling are removed; a sweep in row
                    order makes each unassigned row a root that takes its
                    strong neighbours (even ones already taken) and,
                    provisionally, their unassigned strong neighbours; emptied
                    aggregates are dropped (plain_aggregates.hpp:137-206)
  prolongation      piecewise constant per component, nv*i+k -> nv*agg(i)+k,
                    no entry for removed rows (tentative_prolongation.hpp:195-205,
                    pointwise_aggregates.hpp:115-129)
  coarse operator   P^T A P scaled by 1/over_interp, over_interp = 1.5 (a float)
                    (aggregation.hpp:95-97, 188-196)
  hierarchy         new levels while the scalar rows exceed coarse_enough = 500;
                    the coarsest is solved directly (amg.hpp:418-470,
                    backend/detail/default_direct_solver.hpp:65); a level
                    whose aggregation finds no aggregate ends the hierarchy
                    and is relaxed instead (amg.hpp:436-447, 510-520)
  relaxation        damped Jacobi x += (0.72 / a_rr) (f - A x)
                    (relaxation/damped_jacobi.hpp:60-133); SPAI0
                    x += (a_rr / sum_c a_rc^2) (f - A x) (relaxation/spai0.hpp:60-110);
                    damped block Jacobi on the diagonal blocks is the GPU's
                    addition
  cycle             V-cycle from x = 0, one pre- and one post-sweep
                    (amg.hpp:287-304, 498-541)

The GPU picks its aggregates with a parallel rule (roots pairwise at least
three strong links apart, chosen in rounds by hashed priority), not the row
sweep; everything after the aggregate ids is built by ``hierarchy`` from
whichever ids it is given, so the GPU hierarchy is checked level by level
against the same ids.  ``gmres`` is a textbook right-preconditioned GMRES(m)
(Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., alg. 9.5) used
only to compare iteration counts.
"""
import numpy as np
import scipy.sparse as sp

EPS_STRONG = 0.08
OVER_INTERP = 1.5
DAMPING = 0.72
COARSE_ENOUGH = 500

REMOVED = -1


def _f32(x):
    return float(np.float32(x))


def eps_squared(eps=EPS_STRONG):
    """AMGCL squares the float parameter in float (plain_aggregates.hpp:122)."""
    e = np.float32(eps)
    return float(e * e)


def galerkin_scale(over=OVER_INTERP):
    """1 / over_interp evaluated in float, as aggregation.hpp:194 does."""
    return float(np.float32(1.0) / np.float32(over))


class BlockCsr:
    """Block CSR with nv x nv blocks (val[nnzb][nv][nv])."""

    def __init__(self, rp, col, val):
        self.rp = np.asarray(rp, np.int64)
        self.col = np.asarray(col, np.int64)
        self.val = np.asarray(val, np.float64)
        self.n = len(self.rp) - 1
        self.nv = self.val.shape[1]

    @staticmethod
    def from_scalar_csr(rp, col, val, nv):
        """The reference's scalar CSR, whose nv rows of a cell share one block column
        pattern (CoefficientMatrix.cpp:114-164), as block CSR in the same order."""
        rp, col, val = np.asarray(rp), np.asarray(col), np.asarray(val)
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
        return BlockCsr(brp, bcol, bval)

    def scalar(self):
        return sp.bsr_matrix((self.val, self.col, self.rp), shape=(self.n * self.nv,) * 2).tocsr()

    def rows(self):
        return np.repeat(np.arange(self.n), np.diff(self.rp))


def pointwise(A):
    """Largest |entry| of every block and of each diagonal block."""
    s = np.abs(A.val).reshape(len(A.col), -1).max(axis=1)
    r = A.rows()
    dia = np.zeros(A.n)
    d = A.col == r
    dia[r[d]] = s[d]
    return s, dia


def strong_connections(A, eps=EPS_STRONG):
    """Per stored block: strong when eps^2 d_i d_j < s_ij^2, never the diagonal."""
    s, dia = pointwise(A)
    r = A.rows()
    e2 = eps_squared(eps)
    return (A.col != r) & ((e2 * dia[r]) * dia[A.col] < s * s)


def plain_aggregates(A, strong):
    """The row sweep of plain_aggregates.hpp:137-206.  Returns (count, id), id -1
    for removed rows."""
    n = A.n
    UNDEF, REM = -1, -2
    idv = np.full(n, REM, dtype=np.int64)
    for i in range(n):
        if strong[A.rp[i]:A.rp[i + 1]].any():
            idv[i] = UNDEF
    count = 0
    for i in range(n):
        if idv[i] != UNDEF:
            continue
        cur = count
        count += 1
        idv[i] = cur
        neib = []
        for p in range(A.rp[i], A.rp[i + 1]):
            c = A.col[p]
            if strong[p] and idv[c] != REM:
                idv[c] = cur
                neib.append(c)
        for c in neib:
            for p in range(A.rp[c], A.rp[c + 1]):
                cc = A.col[p]
                if strong[p] and idv[cc] == UNDEF:
                    idv[cc] = cur
    # renumber past aggregates that lost every member
    used = np.zeros(count, dtype=np.int64)
    used[idv[idv >= 0]] = 1
    newid = np.cumsum(used) - 1
    out = np.where(idv >= 0, newid[np.maximum(idv, 0)], REMOVED)
    return int(used.sum()), out


def prolongation(agg, naggr, nv):
    """Scalar P: row nv*i + k has a one in column nv*agg(i) + k (none if removed)."""
    n = len(agg)
    rows = np.arange(n * nv)
    a = np.repeat(agg, nv)
    keep = a >= 0
    colp = (a * nv + np.tile(np.arange(nv), n))[keep]
    return sp.csr_matrix((np.ones(keep.sum()), (rows[keep], colp)), shape=(n * nv, naggr * nv))


def galerkin(A, agg, naggr, over=OVER_INTERP):
    """Coarse block CSR: block (I, J) = scale * sum of blocks (i, j) with
    agg(i) = I, agg(j) = J (the block form of P^T A P for piecewise-constant P)."""
    r = A.rows()
    I, J = agg[r], agg[A.col]
    keep = (I >= 0) & (J >= 0)
    key = I[keep] * naggr + J[keep]
    uk, inv = np.unique(key, return_inverse=True)
    val = np.zeros((len(uk), A.nv, A.nv))
    np.add.at(val, inv, A.val[keep])
    val *= galerkin_scale(over)
    rows = uk // naggr
    rp = np.zeros(naggr + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    return BlockCsr(np.cumsum(rp), uk % naggr, val)


def smoother_diagonal(A, kind="jacobi", damping=DAMPING):
    """Per scalar row the factor of x += w (f - A x): damping / a_rr (damped
    Jacobi) or a_rr / sum_c a_rc^2 (SPAI0)."""
    S = A.scalar()
    d = S.diagonal()
    if kind == "jacobi":
        return damping * (1.0 / d)
    return d / np.asarray(S.multiply(S).sum(axis=1)).ravel()


def smoother(A, kind="jacobi", damping=DAMPING):
    """The correction W t of one sweep x += W (f - A x).  "block" is damped
    block Jacobi on the nv x nv diagonal blocks, W = damping D^-1 (a relaxation
    the GPU offers beside the reference's two scalar ones)."""
    if kind == "block":
        r = A.rows()
        d = A.col == r
        D = np.zeros((A.n, A.nv, A.nv))
        D[r[d]] = A.val[d]
        W = damping * np.linalg.inv(D)
        return lambda t: np.einsum("nij,nj->ni", W, t.reshape(A.n, A.nv)).ravel()
    w = smoother_diagonal(A, kind, damping)
    return lambda t: w * t


class Level:
    def __init__(self, A, relax):
        self.A, self.S, self.relax = A, A.scalar(), relax
        self.P = None
        self.agg = None


def hierarchy(A, aggregates=None, kind="jacobi", coarse_enough=COARSE_ENOUGH, eps=EPS_STRONG,
              over=OVER_INTERP, damping=DAMPING, max_levels=20):
    """Levels of amg.hpp:418-470.  ``aggregates(level, A)`` returns (count, id) for
    a level; the default is the reference's row sweep.  Returns (levels, inverse
    of the coarsest matrix or None when the coarsest is relaxed instead)."""
    levels = []
    direct = True
    while A.n * A.nv > coarse_enough and len(levels) < max_levels:
        lev = Level(A, smoother(A, kind, damping))
        levels.append(lev)
        if aggregates is None:
            naggr, agg = plain_aggregates(A, strong_connections(A, eps))
        else:
            naggr, agg = aggregates(len(levels) - 1, A)
        if naggr == 0:
            direct = False
            break
        lev.agg, lev.P = agg, prolongation(agg, naggr, A.nv)
        A = galerkin(A, agg, naggr, over)
    if len(levels) == max_levels or not direct:
        return levels, None
    coarse = Level(A, smoother(A, kind, damping))
    levels.append(coarse)
    return levels, np.linalg.inv(A.scalar().toarray())


def vcycle(levels, inv, f, l=0):
    """One V-cycle from x = 0 (amg.hpp:287-304, 498-541)."""
    lev = levels[l]
    if l == len(levels) - 1:
        if inv is not None:
            return inv @ f
        x = lev.relax(f)
        return x + lev.relax(f - lev.S @ x)
    x = lev.relax(f)                       # pre-sweep from x = 0
    t = f - lev.S @ x
    u = vcycle(levels, inv, lev.P.T @ t, l + 1)
    x = x + lev.P @ u
    return x + lev.relax(f - lev.S @ x)    # post-sweep


def gmres(A, M, b, m=30, rtol=1e-8, maxiter=1000):
    """Right-preconditioned restarted GMRES(m) with modified Gram-Schmidt and
    Givens rotations, x0 = 0.  Returns (x, iterations, relative residual)."""
    n = len(b)
    x = np.zeros(n)
    bn = np.linalg.norm(b)
    if bn == 0:
        return x, 0, 0.0
    it = 0
    r = b.copy()
    while it < maxiter:
        beta = np.linalg.norm(r)
        if beta <= rtol * bn:
            break
        V = np.zeros((m + 1, n))
        Z = np.zeros((m, n))
        H = np.zeros((m + 1, m))
        V[0] = r / beta
        g = np.zeros(m + 1)
        g[0] = beta
        cs, sn = np.zeros(m), np.zeros(m)
        j = 0
        while j < m and it < maxiter:
            Z[j] = M(V[j])
            w = A @ Z[j]
            for q in range(j + 1):
                H[q, j] = V[q] @ w
                w = w - H[q, j] * V[q]
            H[j + 1, j] = np.linalg.norm(w)
            if H[j + 1, j] > 0:
                V[j + 1] = w / H[j + 1, j]
            for q in range(j):
                t = cs[q] * H[q, j] + sn[q] * H[q + 1, j]
                H[q + 1, j] = -sn[q] * H[q, j] + cs[q] * H[q + 1, j]
                H[q, j] = t
            den = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = (H[j, j] / den, H[j + 1, j] / den) if den > 0 else (1.0, 0.0)
            H[j, j], H[j + 1, j] = den, 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            j += 1
            it += 1
            if abs(g[j]) <= rtol * bn:
                break
        y = np.linalg.solve(np.triu(H[:j, :j]), g[:j])
        x = x + Z[:j].T @ y
        r = b - A @ x
    return x, it, np.linalg.norm(b - A @ x) / bn
