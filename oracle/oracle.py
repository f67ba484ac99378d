"""ctypes front end of the CPU oracle (oracle/libbes_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product (bes_amd) never
imports it.  See bes_oracle.h for the conventions.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbes_oracle.so")

GAMMA, GAS_R, HARTEN, VENKAT_K = 1.4, 287.0, 0.05, 5.0
RK4_ALPHA = (0.0833, 0.2069, 0.4265, 1.0)
COMPAT, EXACT = 0, 7


class Gas(C.Structure):
    _fields_ = [("gamma", C.c_double), ("R", C.c_double), ("harten", C.c_double),
                ("venkat_k", C.c_double)]


def default_gas():
    return Gas(GAMMA, GAS_R, HARTEN, VENKAT_K)


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        P = C.c_void_p
        _lib.or_mesh_create.restype = P
        _lib.or_mesh_create.argtypes = [C.c_int, C.c_long, P, C.c_long, P, C.c_long, P, P, C.c_int]
        _lib.or_mesh_destroy.argtypes = [P]
        _lib.or_mesh_nface.restype = C.c_long
        _lib.or_mesh_nface.argtypes = [P]
        _lib.or_mesh_kface.restype = C.c_int
        _lib.or_mesh_kface.argtypes = [P]
        _lib.or_mesh_get_faces.argtypes = [P] * 8
        _lib.or_mesh_get_cells.argtypes = [P] * 9
        _lib.or_flux.argtypes = [C.c_int, C.c_int, C.c_long, P, P, P, P, P]
        _lib.or_jacobian.argtypes = [C.c_int, C.c_int, C.c_long, P, P, P, P, P, P]
        _lib.or_gradient.argtypes = [P, P, P]
        _lib.or_limiter.argtypes = [P, P, P, P, P]
        _lib.or_timestep.argtypes = [P, P, C.c_double, C.c_int, P, P]
        _lib.or_residual.argtypes = [P, P, P, P, P, P, P, P]
        _lib.or_explicit.argtypes = [P, P, C.c_int, C.c_int, P, C.c_double, C.c_int, P, P, P, C.c_int]
        _lib.or_implicit_nnzb.restype = C.c_long
        _lib.or_implicit_nnzb.argtypes = [P]
        _lib.or_implicit_assemble.argtypes = [P, P, C.c_double, C.c_int, P, P, C.c_int, P, P, P, P, P]
        _lib.or_implicit_create.restype = P
        _lib.or_implicit_create.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_int]
        _lib.or_implicit_destroy.argtypes = [P]
        _lib.or_implicit_ncolor.restype = C.c_int
        _lib.or_implicit_ncolor.argtypes = [P]
        _lib.or_implicit_steps.restype = C.c_int
        _lib.or_implicit_steps.argtypes = [P, P, P, C.c_int, P, C.c_double, C.c_double, P, C.c_int, C.c_int,
                                           C.c_double, C.c_double, P, P, P, P]
        _lib.or_implicit_last.argtypes = [P, P, P, P]
        _lib.or_implicit_freeze_limiter.argtypes = [P, P]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def freestream_primitive(mach, ndim, gamma=GAMMA, R=GAS_R):
    """Non-dimensional freestream primitive state (Config.cpp:479-488)."""
    p = np.zeros(ndim + 3)
    p[0] = 1.0
    p[1:1 + ndim] = np.asarray(mach, dtype=np.float64)[:ndim]
    p[ndim + 1] = 1.0 / gamma
    p[ndim + 2] = 1.0 / (gamma * R)
    return p


def conserved_from_primitive(P, ndim, gamma=GAMMA, R=GAS_R):
    """GasPerfect.tpp:181-216 for one state (cell-major arrays broadcast)."""
    P = np.asarray(P, dtype=np.float64)
    cv = R / (gamma - 1.0)
    U = np.zeros(P.shape[:-1] + (ndim + 2,))
    U[..., 0] = P[..., 0]
    q2 = np.zeros(P.shape[:-1])
    for d in range(ndim):
        U[..., 1 + d] = U[..., 0] * P[..., 1 + d]
        q2 = q2 + P[..., 1 + d] * P[..., 1 + d]
    U[..., ndim + 1] = U[..., 0] * (cv * P[..., ndim + 2] + q2 / 2.0)
    return U


class Mesh:
    """Oracle mesh: reference face numbering and metric definitions."""

    def __init__(self, md, geom_mode=COMPAT):
        L = lib()
        self.ndim = md.ndim
        self.nv = md.ndim + 2
        self.k = md.ndim + 1
        self.ncell = md.ncell
        self._keep = [np.ascontiguousarray(md.nodes, dtype=np.float64),
                      np.ascontiguousarray(md.cells, dtype=np.int64),
                      np.ascontiguousarray(md.bfaces, dtype=np.int64),
                      np.ascontiguousarray(md.btype, dtype=np.int32)]
        X, cells, bf, bt = self._keep
        self.h = L.or_mesh_create(md.ndim, X.shape[0], _p(X), cells.shape[0], _p(cells),
                                  bf.shape[0], _p(bf), _p(bt), geom_mode)
        self.nface = L.or_mesh_nface(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mesh_destroy(self.h)
            self.h = None

    def faces(self):
        nf, nd = self.nface, self.ndim
        out = dict(owner=np.zeros(nf, np.int64), nbr=np.zeros(nf, np.int64),
                   btype=np.zeros(nf, np.int32), S=np.zeros((nf, nd)), n=np.zeros((nf, nd)),
                   area=np.zeros(nf), center=np.zeros((nf, nd)))
        lib().or_mesh_get_faces(self.h, *[_p(out[k]) for k in
                                          ("owner", "nbr", "btype", "S", "n", "area", "center")])
        return out

    def cells(self):
        ne, nd, k = self.ncell, self.ndim, self.k
        out = dict(vol=np.zeros(ne), centroid=np.zeros((ne, nd)), cface=np.zeros((ne, k), np.int64),
                   cadj=np.zeros((ne, k), np.int64), w=np.zeros((ne, k)), r=np.zeros((ne, k, nd)),
                   rsel=np.zeros((ne, k), np.int64), sproj=np.zeros((ne, nd)))
        lib().or_mesh_get_cells(self.h, *[_p(out[k_]) for k_ in
                                          ("vol", "centroid", "cface", "cadj", "w", "r", "rsel", "sproj")])
        return out

    # ---- field quantities (U is [ncell][nv]) ----
    def gradient(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        G = np.zeros((self.ncell, self.nv, self.ndim))
        lib().or_gradient(self.h, _p(U), _p(G))
        return G

    def limiter(self, U, G, gas=None):
        gas = gas or default_gas()
        U = np.ascontiguousarray(U, dtype=np.float64)
        G = np.ascontiguousarray(G, dtype=np.float64)
        P = np.zeros((self.ncell, self.nv))
        lib().or_limiter(self.h, _p(U), _p(G), C.byref(gas), _p(P))
        return P

    def timestep(self, U, cfl, local=True, gas=None):
        gas = gas or default_gas()
        U = np.ascontiguousarray(U, dtype=np.float64)
        dt = np.zeros(self.ncell)
        lib().or_timestep(self.h, _p(U), cfl, int(local), C.byref(gas), _p(dt))
        return dt

    def residual(self, U, G, phi, finf, gas=None, with_flux=False):
        gas = gas or default_gas()
        U = np.ascontiguousarray(U, dtype=np.float64)
        R = np.zeros((self.ncell, self.nv))
        F = np.zeros((self.nface, self.nv)) if with_flux else None
        fi = np.ascontiguousarray(finf, dtype=np.float64)
        lib().or_residual(self.h, _p(U), _p(np.ascontiguousarray(G)), _p(np.ascontiguousarray(phi)),
                          _p(fi), C.byref(gas), _p(R), _p(F))
        return (R, F) if with_flux else R

    def explicit(self, U, niter, cfl, finf, local=True, alpha=RK4_ALPHA, gas=None, nthreads=1):
        gas = gas or default_gas()
        U = np.array(U, dtype=np.float64, order="C")
        a = np.asarray(alpha, dtype=np.float64)
        linf = np.zeros((niter, self.nv))
        fi = np.ascontiguousarray(finf, dtype=np.float64)
        lib().or_explicit(self.h, _p(U), niter, len(alpha), _p(a), cfl, int(local), _p(fi),
                          C.byref(gas), _p(linf), nthreads)
        return U, linf

    def implicit_assemble(self, U, cfl, finf, local=True, repeat_nv=True, gas=None):
        gas = gas or default_gas()
        U = np.ascontiguousarray(U, dtype=np.float64)
        nnzb = lib().or_implicit_nnzb(self.h)
        rp = np.zeros(self.ncell + 1, np.int64)
        col = np.zeros(nnzb, np.int64)
        val = np.zeros((nnzb, self.nv, self.nv))
        rhs = np.zeros((self.ncell, self.nv))
        dt = np.zeros(self.ncell)
        fi = np.ascontiguousarray(finf, dtype=np.float64)
        lib().or_implicit_assemble(self.h, _p(U), cfl, int(local), _p(fi), C.byref(gas),
                                   int(repeat_nv), _p(rp), _p(col), _p(val), _p(rhs), _p(dt))
        return rp, col, val, rhs, dt


class Implicit:
    """The implicit pseudo-time loop on the CPU (SolverImplicit.cpp:552-600):
    OpenMP assembly with atomic adds, multicolour block ILU(0)-GMRES(restart)
    to the relative tolerance, adaptive Courant number.  Keeps the Courant
    number and the previous Linf(b) between calls, like the device's
    bes_implicit_run / bes_implicit_continue."""

    def __init__(self, mesh, restart=30, maxiter=300, rtol=5e-2, nthreads=0):
        self.mesh = mesh
        self.h = lib().or_implicit_create(mesh.h, restart, maxiter, rtol, nthreads)
        self.ncolor = lib().or_implicit_ncolor(self.h)
        self.cfl = None
        self.lb_old = np.zeros(mesh.nv)
        self.r0 = -1.0

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_implicit_destroy(self.h)
            self.h = None

    def start(self, cfl_min, cfl_max, exponents, local=True, repeat_nv=False, stop_drop=0.0):
        """A new loop: the first step runs at cfl_min (SolverImplicit.cpp:347-356)."""
        self.cfg = (float(cfl_min), float(cfl_max), np.asarray(exponents, np.float64), int(local),
                    int(repeat_nv), float(stop_drop))
        self.cfl = C.c_double(cfl_min)
        self.lb_old = np.zeros(self.mesh.nv)
        self.r0 = -1.0

    def resume(self, cfl, lb_old, r0):
        """Continue a loop at Courant number cfl with the previous step's Linf(b)."""
        self.cfl = C.c_double(cfl)
        self.lb_old = np.ascontiguousarray(lb_old, dtype=np.float64).copy()
        self.r0 = float(r0)

    def steps(self, U, nsteps, finf, gas=None):
        """Up to nsteps steps in place on U ([ncell][nv] float64, C order); the
        history rows as bes_implicit_run's."""
        gas = gas or default_gas()
        cmin, cmax, ex, local, rep, drop = self.cfg
        assert U.dtype == np.float64 and U.flags.c_contiguous
        nv = self.mesh.nv
        hist = np.zeros((nsteps, 2 * nv + 3))
        fi = np.ascontiguousarray(finf, dtype=np.float64)
        n = lib().or_implicit_steps(self.h, self.mesh.h, _p(U), nsteps, C.byref(self.cfl), cmin, cmax, _p(ex),
                                    local, rep, drop, self.r0, _p(fi), C.byref(gas), _p(self.lb_old), _p(hist))
        if self.r0 <= 0 and n > 0:
            self.r0 = float(hist[0, nv])
        return hist[:n]

    def freeze_limiter(self, phi):
        """Use the limiter values phi [ncell][nv] in every later assembly
        (None: recompute the limiter again)."""
        self._phi = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
        lib().or_implicit_freeze_limiter(self.h, _p(self._phi))

    def last_system(self):
        nv = self.mesh.nv
        nnzb = lib().or_implicit_nnzb(self.mesh.h)
        val = np.zeros((nnzb, nv, nv))
        rhs = np.zeros((self.mesh.ncell, nv))
        x = np.zeros((self.mesh.ncell, nv))
        lib().or_implicit_last(self.h, _p(val), _p(rhs), _p(x))
        return val, rhs, x


def flux(ndim, kind, UL, UR, n, gas=None):
    """Batch face flux: UR rows have nv+1 entries (primitive freestream for kind 3)."""
    gas = gas or default_gas()
    nv = ndim + 2
    UL = np.ascontiguousarray(UL, dtype=np.float64).reshape(-1, nv)
    URp = np.zeros((UL.shape[0], nv + 1))
    UR = np.asarray(UR, dtype=np.float64).reshape(UL.shape[0], -1)
    URp[:, :UR.shape[1]] = UR
    n = np.ascontiguousarray(n, dtype=np.float64).reshape(-1, ndim)
    F = np.zeros((UL.shape[0], nv))
    lib().or_flux(ndim, kind, UL.shape[0], _p(UL), _p(URp), _p(n), C.byref(gas), _p(F))
    return F


def jacobian(ndim, kind, UL, UR, n, gas=None):
    gas = gas or default_gas()
    nv = ndim + 2
    UL = np.ascontiguousarray(UL, dtype=np.float64).reshape(-1, nv)
    URp = np.zeros((UL.shape[0], nv + 1))
    UR = np.asarray(UR, dtype=np.float64).reshape(UL.shape[0], -1)
    URp[:, :UR.shape[1]] = UR
    n = np.ascontiguousarray(n, dtype=np.float64).reshape(-1, ndim)
    JL = np.zeros((UL.shape[0], nv, nv))
    JR = np.zeros((UL.shape[0], nv, nv))
    lib().or_jacobian(ndim, kind, UL.shape[0], _p(UL), _p(URp), _p(n), C.byref(gas), _p(JL), _p(JR))
    return JL, JR


def perturbed_state(ncell, ndim, mach, seed, amp=0.05):
    """Freestream conserved state with seeded +-amp relative perturbations."""
    rng = np.random.default_rng(seed)
    P = freestream_primitive(mach, ndim)
    U = np.tile(conserved_from_primitive(P, ndim), (ncell, 1))
    U *= 1.0 + rng.uniform(-amp, amp, size=U.shape)
    return U
