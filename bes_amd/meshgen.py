"""Synthetic unstructured meshes for the BES solver core, plus a Gmsh MSH2 writer.

Every generator returns a :class:`MeshData`: node coordinates, volume elements
(CCW triangles in 2D, positively oriented tetrahedra in 3D) and boundary faces
with a boundary type, listed in the order a Gmsh MSH2 file would hold them
(boundary elements first, then volume elements).  That order matters: the
reference numbers faces "boundary faces in file order, then internal faces in
discovery order" (Mesh.cpp:236-534), and the solver core reproduces it.

The MSH2 subset written here is the one the reference reader accepts
(Mesh.cpp:95-558): ``$PhysicalNames`` before ``$Nodes``, exactly two tags per
element with the physical tag first, element types 1 (line), 2 (triangle) and
4 (tetrahedron) only, contiguous node ids from 1, single-space separators.
"""
from __future__ import annotations

import dataclasses
import numpy as np

# Boundary type codes, matching the reference enum BoundaryType
# (Options.hpp:34-35): INTERNAL=1, SLIP_WALL=2, FREESTREAM=3.
INTERNAL = 1
SLIP_WALL = 2
FREESTREAM = 3
BOUNDARY_NAMES = {SLIP_WALL: "SLIP_WALL", FREESTREAM: "FREESTREAM"}


@dataclasses.dataclass
class MeshData:
    ndim: int
    nodes: np.ndarray        # [nnode, ndim] float64
    cells: np.ndarray        # [ncell, ndim+1] int64, 0-based node ids
    bfaces: np.ndarray       # [nbface, ndim] int64, 0-based node ids
    btype: np.ndarray        # [nbface] int32 boundary type code

    @property
    def ncell(self) -> int:
        return int(self.cells.shape[0])

    @property
    def nnode(self) -> int:
        return int(self.nodes.shape[0])


# --------------------------------------------------------------------------
# helpers
# --------------------------------------------------------------------------

def _orient_tris(nodes, tris):
    """Return triangles in counter-clockwise order (reference Q10 for 2D)."""
    p = nodes[tris]
    cross = ((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
             - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    tris = tris.copy()
    neg = cross < 0
    tris[neg, 1], tris[neg, 2] = tris[neg, 2].copy(), tris[neg, 1].copy()
    return tris


def tet_signed_volume6(nodes, tets):
    p = nodes[tets]
    a = p[:, 1] - p[:, 0]
    b = p[:, 2] - p[:, 0]
    c = p[:, 3] - p[:, 0]
    return np.einsum("ij,ij->i", a, np.cross(b, c))


def _orient_tets(nodes, tets):
    """Return tetrahedra with positive orientation (reference Q10 for 3D)."""
    v = tet_signed_volume6(nodes, tets)
    tets = tets.copy()
    neg = v < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    return tets


def boundary_faces_of(cells: np.ndarray, ndim: int) -> np.ndarray:
    """Faces (as node tuples) that belong to exactly one cell."""
    if ndim == 2:
        loc = [(0, 1), (1, 2), (2, 0)]
    else:
        loc = [(0, 2, 1), (1, 2, 3), (0, 3, 2), (0, 1, 3)]
    faces = np.concatenate([cells[:, list(l)] for l in loc], axis=0)
    key = np.sort(faces, axis=1).astype(np.int64)
    if int(cells.max()) < (1 << 21):
        packed = key[:, 0]
        for q in range(1, key.shape[1]):
            packed = (packed << 21) | key[:, q]
        _, inv, cnt = np.unique(packed, return_inverse=True, return_counts=True)
    else:
        _, inv, cnt = np.unique(key, axis=0, return_inverse=True, return_counts=True)
    inv = inv.reshape(-1)
    once = cnt[inv] == 1
    return faces[once]


# --------------------------------------------------------------------------
# 2D channel with a circular-arc bump (config: 200x100 quads split in 40k tris)
# --------------------------------------------------------------------------

def bump_lower_wall(x, thickness=0.1, x0=1.0, x1=2.0):
    """Circular arc of the given thickness on x in [x0, x1], zero elsewhere."""
    half = 0.5 * (x1 - x0)
    radius = (half * half + thickness * thickness) / (2.0 * thickness)
    xc = 0.5 * (x0 + x1)
    y = np.sqrt(np.maximum(radius * radius - (x - xc) ** 2, 0.0)) - (radius - thickness)
    return np.where((x > x0) & (x < x1), np.maximum(y, 0.0), 0.0)


def bump_channel(nx=200, ny=100, seed=20261202, jitter=0.1, length=3.0, height=1.0):
    """2D channel [0,length]x[0,height] with a 10%-thick circular-arc bump.

    (nx+1)x(ny+1) nodes, algebraic spacing from the lower wall to the top,
    each quad split on a seeded random diagonal, interior nodes jittered by
    ``jitter`` times the local spacing.  Lower and upper walls are slip walls,
    inlet and outlet are freestream boundaries.
    """
    rng_diag = np.random.default_rng(seed)
    rng_jit = np.random.default_rng(seed + 1)
    xs = np.linspace(0.0, length, nx + 1)
    yb = bump_lower_wall(xs)
    eta = np.linspace(0.0, 1.0, ny + 1)
    X = np.repeat(xs[:, None], ny + 1, axis=1)
    Y = yb[:, None] + eta[None, :] * (height - yb[:, None])
    dx = length / nx
    dy = (height - yb[:, None]) / ny * np.ones((1, ny + 1))
    jx = rng_jit.uniform(-jitter, jitter, size=X.shape) * dx
    jy = rng_jit.uniform(-jitter, jitter, size=X.shape) * dy
    interior = np.zeros_like(X, dtype=bool)
    interior[1:-1, 1:-1] = True
    X = X + np.where(interior, jx, 0.0)
    Y = Y + np.where(interior, jy, 0.0)
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)

    def nid(i, j):
        return i * (ny + 1) + j

    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    I = I.ravel()
    J = J.ravel()
    a = nid(I, J)
    b = nid(I + 1, J)
    c = nid(I + 1, J + 1)
    d = nid(I, J + 1)
    diag = rng_diag.random(I.shape[0]) < 0.5
    t1 = np.where(diag[:, None], np.stack([a, b, c], 1), np.stack([a, b, d], 1))
    t2 = np.where(diag[:, None], np.stack([a, c, d], 1), np.stack([b, c, d], 1))
    tris = np.stack([t1, t2], axis=1).reshape(-1, 3)
    tris = _orient_tris(nodes, tris)

    # boundary edges in a fixed order: bottom, right, top, left
    bottom = np.stack([nid(np.arange(nx), 0), nid(np.arange(1, nx + 1), 0)], 1)
    right = np.stack([nid(nx, np.arange(ny)), nid(nx, np.arange(1, ny + 1))], 1)
    top = np.stack([nid(np.arange(nx, 0, -1), ny), nid(np.arange(nx - 1, -1, -1), ny)], 1)
    left = np.stack([nid(0, np.arange(ny, 0, -1)), nid(0, np.arange(ny - 1, -1, -1))], 1)
    bfaces = np.concatenate([bottom, right, top, left], axis=0)
    btype = np.concatenate([
        np.full(nx, SLIP_WALL), np.full(ny, FREESTREAM),
        np.full(nx, SLIP_WALL), np.full(ny, FREESTREAM)]).astype(np.int32)
    return MeshData(2, nodes, tris.astype(np.int64), bfaces.astype(np.int64), btype)


# --------------------------------------------------------------------------
# 3D box of Kuhn-split hexahedra (tests and synthetic throughput meshes)
# --------------------------------------------------------------------------

_KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def kuhn_tets(ni, nj, nk, node_id):
    """Split each hexahedron of an ni x nj x nk lattice into 6 Kuhn tetrahedra.

    ``node_id(i, j, k)`` maps lattice corners (arrays) to node ids.  The split
    along the (0,0,0)-(1,1,1) diagonal of every hex is conforming across hex
    faces, so the result is a valid tetrahedral mesh.
    """
    I, J, K = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    I = I.ravel(); J = J.ravel(); K = K.ravel()
    base = np.stack([I, J, K], 1)
    tets = []
    for perm in _KUHN:
        p = base.copy()
        verts = [node_id(p[:, 0], p[:, 1], p[:, 2])]
        for ax in perm:
            p = p.copy()
            p[:, ax] += 1
            verts.append(node_id(p[:, 0], p[:, 1], p[:, 2]))
        tets.append(np.stack(verts, 1))
    return np.stack(tets, 1).reshape(-1, 4)


def box_tets(ni=4, nj=3, nk=3, seed=7, jitter=0.15, extent=(1.0, 1.0, 1.0),
             slip_faces=("zmin",)):
    """Tetrahedral box mesh with jittered interior nodes.

    Faces named in ``slip_faces`` (xmin/xmax/ymin/ymax/zmin/zmax) are slip
    walls, the rest are freestream boundaries.
    """
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, extent[0], ni + 1)
    ys = np.linspace(0.0, extent[1], nj + 1)
    zs = np.linspace(0.0, extent[2], nk + 1)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    P = np.stack([X, Y, Z], -1)
    h = np.array([extent[0] / ni, extent[1] / nj, extent[2] / nk])
    interior = np.zeros(X.shape, dtype=bool)
    interior[1:-1, 1:-1, 1:-1] = True
    P = P + np.where(interior[..., None], rng.uniform(-jitter, jitter, P.shape) * h, 0.0)
    nodes = P.reshape(-1, 3)

    def node_id(i, j, k):
        return (i * (nj + 1) + j) * (nk + 1) + k

    tets = _orient_tets(nodes, kuhn_tets(ni, nj, nk, node_id))
    bf = boundary_faces_of(tets, 3)
    c = nodes[bf].mean(axis=1)
    eps = 1e-9
    names = {
        "xmin": np.abs(c[:, 0]) < eps, "xmax": np.abs(c[:, 0] - extent[0]) < eps,
        "ymin": np.abs(c[:, 1]) < eps, "ymax": np.abs(c[:, 1] - extent[1]) < eps,
        "zmin": np.abs(c[:, 2]) < eps, "zmax": np.abs(c[:, 2] - extent[2]) < eps,
    }
    btype = np.full(bf.shape[0], FREESTREAM, dtype=np.int32)
    for nm in slip_faces:
        btype[names[nm]] = SLIP_WALL
    # shuffle face vertex order so the reader has to orient normals itself
    rot = rng.integers(0, 3, size=bf.shape[0])
    bf = np.take_along_axis(bf, (np.arange(3)[None, :] - rot[:, None]) % 3, axis=1)
    return MeshData(3, nodes, tets.astype(np.int64), bf.astype(np.int64), btype)


def tri_patch(ni=4, nj=3, seed=3, jitter=0.2, slip=("bottom",)):
    """Small 2D jittered triangle mesh on the unit square (tests)."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, 1.0, ni + 1)
    ys = np.linspace(0.0, 1.0, nj + 1)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    interior = np.zeros(X.shape, dtype=bool)
    interior[1:-1, 1:-1] = True
    X = X + np.where(interior, rng.uniform(-jitter, jitter, X.shape) / ni, 0.0)
    Y = Y + np.where(interior, rng.uniform(-jitter, jitter, Y.shape) / nj, 0.0)
    nodes = np.stack([X.ravel(), Y.ravel()], 1)

    def nid(i, j):
        return i * (nj + 1) + j

    I, J = np.meshgrid(np.arange(ni), np.arange(nj), indexing="ij")
    I = I.ravel(); J = J.ravel()
    a, b, c, d = nid(I, J), nid(I + 1, J), nid(I + 1, J + 1), nid(I, J + 1)
    diag = rng.random(I.shape[0]) < 0.5
    t1 = np.where(diag[:, None], np.stack([a, b, c], 1), np.stack([a, b, d], 1))
    t2 = np.where(diag[:, None], np.stack([a, c, d], 1), np.stack([b, c, d], 1))
    tris = _orient_tris(nodes, np.stack([t1, t2], 1).reshape(-1, 3))
    bf = boundary_faces_of(tris, 2)
    cen = nodes[bf].mean(axis=1)
    eps = 1e-9
    names = {"bottom": np.abs(cen[:, 1]) < eps, "top": np.abs(cen[:, 1] - 1) < eps,
             "left": np.abs(cen[:, 0]) < eps, "right": np.abs(cen[:, 0] - 1) < eps}
    btype = np.full(bf.shape[0], FREESTREAM, dtype=np.int32)
    for nm in slip:
        btype[names[nm]] = SLIP_WALL
    flip = rng.random(bf.shape[0]) < 0.5
    bf = np.where(flip[:, None], bf[:, ::-1], bf)
    return MeshData(2, nodes, tris.astype(np.int64), bf.astype(np.int64), btype)


# --------------------------------------------------------------------------
# 3D swept tapered wing in a half-span box
# --------------------------------------------------------------------------

NACA_CLOSED_TE = (0.2969, -0.1260, -0.3516, 0.2843, -0.1036)


def naca_half_thickness(xi, t=0.10, coeff=NACA_CLOSED_TE):
    """Half thickness of the 4-digit section per unit chord (closed trailing edge)."""
    xi = np.clip(xi, 0.0, 1.0)
    return 5.0 * t * (coeff[0] * np.sqrt(xi) + coeff[1] * xi + coeff[2] * xi ** 2
                      + coeff[3] * xi ** 3 + coeff[4] * xi ** 4)


@dataclasses.dataclass
class WingSpec:
    root_chord: float = 1.0
    taper: float = 0.56
    aspect_ratio: float = 3.8
    sweep_le_deg: float = 30.0
    thickness: float = 0.10
    farfield: float = 10.0        # chords
    tip_cap: float = 0.04         # fraction of the half span closed by an elliptic taper

    @property
    def half_span(self):
        return self.aspect_ratio * self.root_chord * (1.0 + self.taper) / 4.0

    @property
    def area_half(self):
        return 0.5 * self.half_span * self.root_chord * (1.0 + self.taper)

    @property
    def mac(self):
        lam = self.taper
        return 2.0 / 3.0 * self.root_chord * (1 + lam + lam * lam) / (1 + lam)


def _stretch(n, h0, length):
    """n cell sizes starting at h0 growing geometrically to cover length."""
    if n <= 0:
        return np.zeros(0)
    lo, hi = 1.0, 2.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        tot = h0 * (r ** n - 1) / (r - 1) if abs(r - 1) > 1e-12 else h0 * n
        if tot > length:
            hi = r
        else:
            lo = r
    r = 0.5 * (lo + hi)
    s = h0 * np.cumsum(r ** np.arange(n))
    return s * (length / s[-1])


def swept_wing(ni=96, nj=48, nk=72, seed=20261202, spec: WingSpec = None, jitter=0.08):
    """Tetrahedral mesh around a swept tapered wing in a half-span box.

    The lattice has ni x nj x nk hexahedra (6 tetrahedra each).  The wing is a
    slit in the lattice: on the plane k = nk/2, nodes inside the planform
    (chordwise 50% of the i range, spanwise ~60% of the j range) are
    duplicated into an upper and a lower copy pushed onto the thickness
    surface; the leading edge, trailing edge and tip lines stay shared, where
    the thickness vanishes (closed trailing edge, elliptic tip taper).  Node
    columns between the surface and the far field are compressed so every
    hexahedron stays valid; cells cluster geometrically toward the wing.
    Interior nodes away from the wing get a seeded jitter.  Boundaries: the
    root plane y = 0, the outboard side plane and the wing are slip walls; the
    upstream, downstream, bottom and top planes are freestream.  The first
    vertical spacing grows with the in-plane distance from the planform and
    the chordwise clustering relaxes away from the wing plane and outboard of
    the tip, so far-field cells stay moderately shaped.  x is chordwise, y
    spanwise, z up.
    """
    spec = spec or WingSpec()
    rng = np.random.default_rng(seed)
    c_r, s_half = spec.root_chord, spec.half_span
    tan_sw = np.tan(np.radians(spec.sweep_le_deg))
    ff = spec.farfield * c_r
    iLE = ni // 4
    iTE = iLE + ni // 2
    jt = int(round(nj * 0.62))
    kS = nk // 2
    nchord = iTE - iLE

    # spanwise stations: uniform-ish on the wing, clustered toward the tip,
    # then geometric out to the far field
    sj = np.linspace(0.0, 1.0, jt + 1)
    y_wing = s_half * np.sin(0.5 * np.pi * sj) ** 1.0 * 0.5 + s_half * sj * 0.5
    hy = y_wing[-1] - y_wing[-2]
    y_out = s_half + _stretch(nj - jt, hy, ff)
    Y = np.concatenate([y_wing, y_out])

    def x_le(y):
        return np.minimum(y, s_half) * tan_sw

    def chord(y):
        return c_r * (1.0 - (1.0 - spec.taper) * np.minimum(y, s_half) / s_half)

    # chordwise parameter on the wing (cosine clustering at LE and TE)
    xi = 0.5 * (1.0 - np.cos(np.pi * np.linspace(0.0, 1.0, nchord + 1)))
    # vertical distribution from the slit plane: first spacing hz0 over the
    # planform, growing with the in-plane distance d from it, so cells far
    # from the wing are not flattened by the clustering near z = 0
    hz0 = 0.004 * c_r
    hz_cap = 0.5 * ff / max(kS, 1)

    NI, NJ, NK = ni + 1, nj + 1, nk + 1
    X = np.zeros((NI, NJ))       # chordwise positions on the wing plane (clustered)
    Xs = np.zeros((NI, NJ))      # the same without LE/TE clustering, used away from it
    Zs_up = np.zeros((NI, NJ))   # surface z of the upper copy (0 off the wing)
    xu = np.linspace(0.0, 1.0, nchord + 1)
    for j in range(NJ):
        y = Y[j]
        xl, c = x_le(y), chord(y)
        xw = xl + c * xi
        h_le = c * (xi[1] - xi[0]) * 4.0
        h_te = c * (xi[-1] - xi[-2]) * 4.0
        x_up = xl - _stretch(iLE, h_le, ff + xl)[::-1]
        x_dn = xl + c + _stretch(ni - iTE, h_te, ff + c_r - (xl + c))
        X[:, j] = np.concatenate([x_up, xw, x_dn])
        hu = c / nchord
        Xs[:, j] = np.concatenate([xl - _stretch(iLE, hu, ff + xl)[::-1], xl + c * xu,
                                   xl + c + _stretch(ni - iTE, hu, ff + c_r - (xl + c))])
        # outboard of the tip the LE/TE clustering relaxes over half a root chord
        g = np.clip((y - s_half) / (0.5 * c_r), 0.0, 1.0)
        g = g * g * (3.0 - 2.0 * g)
        X[:, j] = (1.0 - g) * X[:, j] + g * Xs[:, j]
        if j <= jt:
            yc = spec.tip_cap * s_half
            tau = 1.0 if y <= s_half - yc else np.sqrt(max(0.0, 1.0 - ((y - (s_half - yc)) / yc) ** 2))
            if j == jt:
                tau = 0.0
            Zs_up[iLE:iTE + 1, j] = c * naca_half_thickness(xi, spec.thickness) * tau
    Zs_up[iLE, :] = 0.0
    Zs_up[iTE, :] = 0.0

    # per-column z levels (nk+1 values, 0 at k = kS)
    Zt = np.zeros((NI, NJ, NK))
    cache = {}
    for j in range(NJ):
        y = Y[j]
        xl, c = x_le(y), chord(y)
        dy = max(y - s_half, 0.0)
        for i in range(NI):
            x = X[i, j]
            dx = max(xl - x, 0.0, x - (xl + c))
            d = np.hypot(dx, dy)
            h0 = min(hz0 * (1.0 + d / (0.05 * c_r)), hz_cap)
            key = round(h0, 12)
            if key not in cache:
                z_up = _stretch(nk - kS, h0, ff)
                z_dn = _stretch(kS, h0, ff)
                cache[key] = np.concatenate([-z_dn[::-1], [0.0], z_up])
            Zt[i, j] = cache[key]

    # node coordinates of the full lattice (upper copies on the slit plane)
    P = np.zeros((NI, NJ, NK, 3))
    # blend from the clustered to the unclustered chordwise distribution over
    # kS/3 layers, starting 3 layers off the wing plane (same weight along a lattice
    # row, so x stays monotone in i)
    tk = np.clip((np.abs(np.arange(NK) - kS) - 3) / max(kS / 3.0, 1.0), 0.0, 1.0)
    beta = tk * tk * (3.0 - 2.0 * tk)
    P[..., 0] = (1.0 - beta)[None, None, :] * X[:, :, None] + beta[None, None, :] * Xs[:, :, None]
    P[..., 1] = Y[None, :, None]
    zs = Zs_up[:, :, None]
    ztop, zbot = ff, ff
    zk = Zt
    upper = zk >= 0
    P[..., 2] = np.where(upper, zs + zk * (1.0 - zs / ztop), -zs + zk * (1.0 - zs / zbot))
    # seeded jitter of interior nodes away from the wing and the boundaries,
    # in lattice-index space: a node moves by a fraction of the vector to its
    # lattice neighbour along each axis, so sheared cells stay valid
    J = rng.uniform(-jitter, jitter, size=P.shape)
    kk = np.arange(NK)[None, None, :]
    near = (np.abs(kk - kS) <= 2) & (np.arange(NI)[:, None, None] >= iLE - 2) & \
           (np.arange(NI)[:, None, None] <= iTE + 2) & (np.arange(NJ)[None, :, None] <= jt + 2)
    inner = np.zeros(P.shape[:3], dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    ii = np.arange(NI)[:, None, None]
    thin = (np.abs(ii - iLE) <= 2) | (np.abs(ii - iTE) <= 2)   # sheared LE/TE columns
    move = inner & ~near & ~thin
    P0 = P.copy()
    for a in range(3):
        fwd = np.zeros_like(P0)
        bwd = np.zeros_like(P0)
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[a] = slice(0, -1)
        sl_hi[a] = slice(1, None)
        fwd[tuple(sl_lo)] = P0[tuple(sl_hi)] - P0[tuple(sl_lo)]
        bwd[tuple(sl_hi)] = P0[tuple(sl_hi)] - P0[tuple(sl_lo)]
        t = J[..., a][..., None]
        step = np.where(t > 0, t * fwd, t * bwd)
        P += np.where(move[..., None], step, 0.0)

    # node ids: lattice nodes, then the lower copies of interior slit nodes
    base = np.arange(NI * NJ * NK).reshape(NI, NJ, NK)
    slit = np.zeros((NI, NJ), dtype=bool)
    slit[iLE + 1:iTE, 0:jt] = True          # interior of the planform (incl. root row j=0)
    nlow = int(slit.sum())
    low_id = np.full((NI, NJ), -1, dtype=np.int64)
    low_id[slit] = NI * NJ * NK + np.arange(nlow)
    lowP = P[:, :, kS, :][slit].copy()
    lowP[:, 2] = -Zs_up[slit]
    nodes = np.concatenate([P.reshape(-1, 3), lowP], axis=0)

    def node_id(i, j, k, below):
        nid = base[i, j, k]
        if below is None:
            return nid
        use_low = below & (k == kS) & slit[i, j]
        return np.where(use_low, low_id[i, j], nid)

    # cells below the slit plane (hex k index = kS-1) use the lower copies
    I, Jj, K = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    I = I.ravel(); Jj = Jj.ravel(); K = K.ravel()
    below_hex = K == kS - 1
    tets = []
    signs = []
    for perm in _KUHN:
        p = np.stack([I, Jj, K], 1)
        verts = [node_id(p[:, 0], p[:, 1], p[:, 2], below_hex)]
        for ax in perm:
            p = p.copy()
            p[:, ax] += 1
            verts.append(node_id(p[:, 0], p[:, 1], p[:, 2], below_hex))
        tets.append(np.stack(verts, 1))
        signs.append(1 if perm in [(0, 1, 2), (1, 2, 0), (2, 0, 1)] else -1)
    tets = np.stack(tets, 1)   # [nhex, 6, 4]
    vol6 = np.stack([tet_signed_volume6(nodes, tets[:, q]) for q in range(6)], 1)
    expect = np.array(signs)[None, :]
    bad = (vol6 * expect <= 0).sum()
    if bad:
        raise ValueError(f"swept_wing: {bad} inverted tetrahedra; coarsen or reduce jitter")
    tets = tets.reshape(-1, 4)
    tets = _orient_tets(nodes, tets)

    bf = boundary_faces_of(tets, 3)
    cen = nodes[bf].mean(axis=1)
    eps = 1e-9
    # far field: upstream, downstream, bottom and top planes.  The outboard
    # side plane runs parallel to the freestream, where the characteristic
    # condition would switch between its inflow and outflow branches on the
    # sign of a vanishing normal velocity; it is a symmetry (slip) plane.
    outer = ((np.abs(cen[:, 0] - nodes[:, 0].min()) < eps) | (np.abs(cen[:, 0] - nodes[:, 0].max()) < eps) |
             (np.abs(cen[:, 2] + ff) < eps) | (np.abs(cen[:, 2] - ff) < eps))
    btype = np.where(outer, FREESTREAM, SLIP_WALL).astype(np.int32)
    md = MeshData(3, nodes, tets.astype(np.int64), bf.astype(np.int64), btype)
    md.wing = dict(spec=spec, iLE=iLE, iTE=iTE, jt=jt, kS=kS)
    return md


def wing_face_mask(md, faces_center, btype):
    """Slip-wall faces on the wing: not the y = 0 symmetry plane and not the
    outboard side plane (the largest y of any slip-wall face)."""
    wall = btype == SLIP_WALL
    y = faces_center[:, 1]
    ymax = y[wall].max() if wall.any() else 0.0
    return wall & (np.abs(y) > 1e-12) & (y < ymax - 1e-9 * max(1.0, abs(ymax)))


def force_coefficients(S, center, btype, owner, U, mach, alpha_deg, spec: WingSpec, gamma=1.4):
    """Lift / drag / side-force coefficients of the wing from cell pressures.

    Pressure force sum_f p_owner S_f over the wing's slip-wall faces (zeroth-
    order face pressure, the value the reference writes to slip_wall_*.csv),
    reference area = half-wing planform area, dynamic pressure 0.5*M^2
    (non-dimensional rho=1, c=1).  Returns (CL, CD, Cp per face, mask).
    """
    m = wing_face_mask(None, center, btype)
    Uc = U[owner[m]]
    rho = Uc[:, 0]
    mom = Uc[:, 1:4]
    p = (gamma - 1.0) * (Uc[:, 4] - 0.5 * (mom * mom).sum(1) / rho)
    F = (p[:, None] * S[m]).sum(0)
    a = np.radians(alpha_deg)
    q = 0.5 * mach * mach
    lift = -F[0] * np.sin(a) + F[2] * np.cos(a)
    drag = F[0] * np.cos(a) + F[2] * np.sin(a)
    cp = (gamma * p - 1.0) / (0.5 * gamma * mach * mach)
    return lift / (q * spec.area_half), drag / (q * spec.area_half), cp, m


# --------------------------------------------------------------------------
# MSH2 writer / reader (the reference's accepted subset)
# --------------------------------------------------------------------------

def write_msh2(path, mesh: MeshData):
    """Write ``mesh`` as Gmsh MSH2 ASCII in the subset the reference reads."""
    nd = mesh.ndim
    face_type = 1 if nd == 2 else 2
    cell_type = 2 if nd == 2 else 4
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", "3",
             f"{nd - 1} {SLIP_WALL} \"SLIP_WALL\"",
             f"{nd - 1} {FREESTREAM} \"FREESTREAM\"",
             f"{nd} 10 \"FLUID\"", "$EndPhysicalNames", "$Nodes", str(mesh.nnode)]
    pts = mesh.nodes
    if nd == 2:
        pts = np.concatenate([pts, np.zeros((pts.shape[0], 1))], axis=1)
    ids = np.arange(1, mesh.nnode + 1)
    lines += [f"{i} {x!r} {y!r} {z!r}" for i, (x, y, z) in zip(ids, pts.tolist())]
    lines += ["$EndNodes", "$Elements", str(mesh.bfaces.shape[0] + mesh.ncell)]
    eid = 1
    for f, t in zip((mesh.bfaces + 1).tolist(), mesh.btype.tolist()):
        lines.append(f"{eid} {face_type} 2 {t} {t} " + " ".join(map(str, f)))
        eid += 1
    for c in (mesh.cells + 1).tolist():
        lines.append(f"{eid} {cell_type} 2 10 10 " + " ".join(map(str, c)))
        eid += 1
    lines += ["$EndElements"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def read_msh2(path) -> MeshData:
    """Read the MSH2 subset written by :func:`write_msh2`."""
    with open(path) as fh:
        text = fh.read().split("\n")
    it = iter(text)
    names = {}
    ndim = 0
    nodes = None
    bf, bt, cells = [], [], []
    for line in it:
        if line == "$PhysicalNames":
            n = int(next(it))
            for _ in range(n):
                d, tag, nm = next(it).split(" ")
                ndim = max(ndim, int(d))
                names[int(tag)] = nm.strip('"')
        elif line == "$Nodes":
            n = int(next(it))
            nodes = np.array([[float(v) for v in next(it).split(" ")[1:1 + 3]] for _ in range(n)])
        elif line == "$Elements":
            n = int(next(it))
            for _ in range(n):
                v = [int(x) for x in next(it).split(" ")]
                et, ntag = v[1], v[2]
                conn = [x - 1 for x in v[3 + ntag:]]
                if (ndim == 2 and et == 1) or (ndim == 3 and et == 2):
                    bf.append(conn)
                    nm = names.get(v[3], "")
                    bt.append(SLIP_WALL if nm == "SLIP_WALL" else FREESTREAM)
                elif (ndim == 2 and et == 2) or (ndim == 3 and et == 4):
                    cells.append(conn)
    return MeshData(ndim, nodes[:, :ndim].copy(), np.array(cells, dtype=np.int64),
                    np.array(bf, dtype=np.int64), np.array(bt, dtype=np.int32))


def write_besi(path, mesh_file, ndim, mach=(0.5, 0.0, 0.0), solver="EXPLICIT",
               cfl=0.8, cfl_max=0.0, exponents=(1.0, 1.0), local_dt=True,
               nsteps=1, output="output"):
    """Write a reference-format ``.besi`` configuration (Config.cpp:116-422)."""
    m = ",".join(repr(float(x)) for x in mach[:ndim])
    lines = [
        "MESH_FORMAT = GMSH", f"MESH_FILE = {mesh_file}", "OUTPUT_FORMAT = VTK",
        f"OUTPUT_FOLDER = {output}", "EQUATION = EULER_NON_DIMENSIONAL",
        f"SOLVER_TYPE = {solver}", "EXPLICIT_SCHEME = MULTI_STEP",
        "MULTI_TIME_STEP_OPTIONS = 4, 0.0833, 0.2069, 0.4265, 1.0",
        f"COURANT_NUM = {cfl!r}", f"ADAPTIVE_COURANT_NUM_MAX = {cfl_max!r}",
        f"ADAPTIVE_COURANT_NUM_EXPONENT = {exponents[0]!r}, {exponents[1]!r}",
        f"LOCAL_TIME_STEP = {'YES' if local_dt else 'NO'}",
        f"NUMBER_OF_TIME_STEP = {nsteps}", "SAVE_INTERVAL = 1000000",
        "FREESTREAM_RAMP = 0", "FREESTREAM_RAMP_START_FRAC = 0.5",
        "GRADIENT_SCHEME = GREEN_GAUSS", "LIMITER_SCHEME = VENKATAKRISHNAN",
        "VENKATAKRISHNAN_FACTOR = 5.0", "FLUX_SCHEME = ROE",
        "HARTEN_CORRECTION_LIMIT = 0.05", "GAS_CONSTANT = 287.0",
        "SPECIFIC_HEAT_RATIO = 1.4", f"FREESTREAM_MACH = {m}",
        "AMG_PRECOND_COARSENER = AGGREGATION", "AMG_PRECOND_SMOOTHER = DAMPED_JACOBI",
        "LINEAR_SOLVER = GMRES", "LINEAR_SOLVER_MAX_ITER = 200",
        "LINEAR_SOLVER_ABS_TOL = 1E-6",
    ]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
