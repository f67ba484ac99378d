"""Python binding of the BES solver core C ABI (include/bes_hip.h).

Loads the in-tree ``bes_amd/libbes_hip.so`` and raises if it is missing:
there is no CPU fallback.  Arrays cross the boundary as plain host buffers in
the reference's conventions (cell-major ``[ncell][nv]`` states in the mesh
file's element order).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbes_hip.so")

OK, ERR_ARG, ERR_HIP, ERR_NAN, ERR_MESH, ERR_SOLVE = range(6)
INTERNAL, SLIP_WALL, FREESTREAM = 1, 2, 3
GEOM_COMPAT, GEOM_EXACT = 0, 7
RK4_ALPHA = (0.0833, 0.2069, 0.4265, 1.0)


class BesError(RuntimeError):
    def __init__(self, status, message, face=-1):
        super().__init__(f"{message} (status {status})")
        self.status = status
        self.face = face


class Gas(C.Structure):
    _fields_ = [("gamma", C.c_double), ("R", C.c_double), ("harten", C.c_double),
                ("venkat_k", C.c_double)]


class MeshDesc(C.Structure):
    _fields_ = [("ndim", C.c_int), ("nnode", C.c_long), ("coords", C.c_void_p),
                ("ncell", C.c_long), ("cells", C.c_void_p), ("nbface", C.c_long),
                ("bfaces", C.c_void_p), ("btype", C.c_void_p)]


def default_gas(gamma=1.4, R=287.0, harten=0.05, venkat_k=5.0):
    return Gas(gamma, R, harten, venkat_k)


def freestream_primitive(mach, ndim, gamma=1.4, R=287.0):
    """Non-dimensional freestream primitive state [rho, u_d, p, T] of
    Config::initializePostMesh (Config.cpp:479-488): rho = 1, u = the Mach
    vector, p = 1/gamma, T = 1/(gamma R), so the sound speed is 1."""
    p = np.zeros(ndim + 3)
    p[0] = 1.0
    p[1:1 + ndim] = np.asarray(mach, dtype=np.float64)[:ndim]
    p[ndim + 1] = 1.0 / gamma
    p[ndim + 2] = 1.0 / (gamma * R)
    return p


def conserved_from_primitive(P, ndim, gamma=1.4, R=287.0):
    """Conserved [rho, rho u_d, rho E] from primitive [rho, u_d, p, T]
    (GasPerfect.tpp:181-216: E = c_v T + q^2/2); leading axes broadcast."""
    P = np.asarray(P, dtype=np.float64)
    cv = R / (gamma - 1.0)
    U = np.empty(P.shape[:-1] + (ndim + 2,))
    U[..., 0] = P[..., 0]
    q2 = np.zeros(P.shape[:-1])
    for d in range(ndim):
        U[..., 1 + d] = P[..., 0] * P[..., 1 + d]
        q2 = q2 + P[..., 1 + d] * P[..., 1 + d]
    U[..., ndim + 1] = P[..., 0] * (cv * P[..., ndim + 2] + q2 / 2.0)
    return U


def wing_freestream(mach, alpha_deg):
    """Mach vector of a wing case: x chordwise, z up, angle of attack alpha."""
    a = np.radians(alpha_deg)
    return (mach * np.cos(a), 0.0, mach * np.sin(a))


_lib = None

_SIGS = {
    "bes_ctx_create": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
    "bes_ctx_create_msh2": [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p],
    "bes_ctx_destroy": [C.c_void_p],
    "bes_ctx_info": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_get_geometry": [C.c_void_p] * 13,
    "bes_set_freestream": [C.c_void_p, C.c_void_p],
    "bes_set_state": [C.c_void_p, C.c_void_p],
    "bes_get_state": [C.c_void_p, C.c_void_p],
    "bes_compute_time_step": [C.c_void_p, C.c_double, C.c_int],
    "bes_get_time_step": [C.c_void_p, C.c_void_p],
    "bes_compute_gradient_limiter": [C.c_void_p],
    "bes_get_gradient": [C.c_void_p, C.c_void_p],
    "bes_get_limiter": [C.c_void_p, C.c_void_p],
    "bes_compute_residual": [C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_explicit_run": [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_double, C.c_int, C.c_void_p],
    "bes_flux_batch": [C.c_int, C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_jacobian_batch": [C.c_int, C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                           C.c_void_p, C.c_void_p],
}
OPTIONAL_SIGS = {
    "bes_implicit_setup": [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double],
    "bes_implicit_assemble": [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_void_p, C.c_void_p],
    "bes_implicit_get_system": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_linear_solve": [C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_get_linear_solution": [C.c_void_p, C.c_void_p],
    "bes_implicit_update": [C.c_void_p],
    "bes_implicit_run": [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_int, C.c_int,
                         C.c_double, C.c_void_p, C.c_void_p],
    "bes_set_profiling": [C.c_void_p, C.c_int],
    "bes_get_profile": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int],
    "bes_synchronize": [C.c_void_p],
    "bes_implicit_set_precond": [C.c_void_p, C.c_int],
    "bes_write_solution": [C.c_void_p, C.c_char_p, C.c_int, C.c_int],
    "bes_write_slip_wall": [C.c_void_p, C.c_char_p, C.c_int],
    "bes_implicit_set_reorth": [C.c_void_p, C.c_double],
    "bes_implicit_stats": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_implicit_set_solver": [C.c_void_p, C.c_int],
    "bes_implicit_set_coloring": [C.c_void_p, C.c_int],
    "bes_autotune": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_implicit_batch_state": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_implicit_ncolors": [C.c_void_p],
    "bes_implicit_continue": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_set_batch": [C.c_void_p, C.c_int],
    "bes_get_batch": [C.c_void_p],
    "bes_set_freestream_batch": [C.c_void_p, C.c_void_p],
    "bes_explicit_graph": [C.c_void_p, C.c_int],
    "bes_explicit_persistent": [C.c_void_p, C.c_int],
    "bes_lp_param_count": [],
    "bes_lp_init": [C.c_void_p, C.c_ulonglong],
    "bes_lp_set_params": [C.c_void_p, C.c_void_p, C.c_int],
    "bes_lp_get_params": [C.c_void_p, C.c_void_p, C.c_int],
    "bes_lp_operator": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_lp_loss_grad": [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p],
    "bes_lp_train": [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_ulonglong, C.c_void_p],
    "bes_lp_fallbacks": [C.c_void_p],
    "bes_lp_set_online": [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int],
    "bes_set_limiter_freeze": [C.c_void_p, C.c_double],
    "bes_limiter_frozen": [C.c_void_p, C.c_void_p],
    "bes_freeze_limiter_values": [C.c_void_p, C.c_void_p],
    "bes_lp_set_rebuild_interval": [C.c_void_p, C.c_int],
    "bes_amg_set_params": [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int],
    "bes_amg_info": [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int],
    "bes_amg_level": [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_amg_apply": [C.c_void_p, C.c_void_p, C.c_void_p],
    "bes_amg_set_aggregation": [C.c_void_p, C.c_int],
}
PROF_KINDS = ["timestep", "gradlim", "residual", "jacobian", "assemble", "spmv", "precond",
              "krylov", "factor", "update", "faceflux", "learned"]


def lib():
    """The loaded C ABI; raises if the HIP library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m bes_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, args in {**_SIGS, **OPTIONAL_SIGS}.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        L.bes_implicit_nnz.argtypes = [C.c_void_p]
        L.bes_lp_online_loss.argtypes = [C.c_void_p]
        L.bes_lp_online_loss.restype = C.c_double
        L.bes_implicit_nnz.restype = C.c_long
        L.bes_last_error_face.argtypes = [C.c_void_p]
        L.bes_last_error_face.restype = C.c_long
        L.bes_last_error_message.restype = C.c_char_p
        L.bes_status_string.restype = C.c_char_p
        L.bes_status_string.argtypes = [C.c_int]
        _lib = L
    return _lib


def exported_symbols():
    """Names declared in include/bes_hip.h (for the export check)."""
    import re
    with open(os.path.join(os.path.dirname(HERE), "include", "bes_hip.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\s*(?:int|long|const char \*|void)\s*\*?\s*(bes_\w+)\s*\(", text, re.M)))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(status, ctx=None):
    if status != OK:
        L = lib()
        msg = L.bes_last_error_message().decode(errors="replace")
        face = L.bes_last_error_face(ctx) if ctx else -1
        raise BesError(status, msg or L.bes_status_string(status).decode(), face)


class Context:
    """A mesh, its geometry and its flow state resident on the GPU."""

    def __init__(self, mesh=None, msh2_path=None, geom_mode=GEOM_COMPAT, gas=None):
        L = lib()
        self.gas = gas or default_gas()
        h = C.c_void_p()
        if msh2_path is not None:
            _check(L.bes_ctx_create_msh2(msh2_path.encode(), geom_mode, C.byref(self.gas), C.byref(h)))
        else:
            self._arrays = [np.ascontiguousarray(mesh.nodes, np.float64),
                            np.ascontiguousarray(mesh.cells, np.int64),
                            np.ascontiguousarray(mesh.bfaces, np.int64),
                            np.ascontiguousarray(mesh.btype, np.int32)]
            X, cells, bf, bt = self._arrays
            desc = MeshDesc(mesh.ndim, X.shape[0], _p(X), cells.shape[0], _p(cells), bf.shape[0],
                            _p(bf), _p(bt))
            _check(L.bes_ctx_create(C.byref(desc), geom_mode, C.byref(self.gas), C.byref(h)))
            self._arrays = None
        self.h = h
        nd, nc, nf = C.c_int(), C.c_long(), C.c_long()
        _check(L.bes_ctx_info(h, C.byref(nd), C.byref(nc), C.byref(nf)))
        self.ndim, self.ncell, self.nface = nd.value, nc.value, nf.value
        self.nv = self.ndim + 2
        self.k = self.ndim + 1
        self.batch = 1

    def _shape(self, *tail):
        # per-condition arrays carry a leading batch axis when batch > 1
        return (self.batch,) + tail if self.batch > 1 else tail

    def close(self):
        if getattr(self, "h", None):
            lib().bes_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- geometry ----
    def geometry(self):
        nf, ne, nd, k = self.nface, self.ncell, self.ndim, self.k
        o = dict(owner=np.zeros(nf, np.int64), nbr=np.zeros(nf, np.int64), btype=np.zeros(nf, np.int32),
                 S=np.zeros((nf, nd)), n=np.zeros((nf, nd)), area=np.zeros(nf), vol=np.zeros(ne),
                 centroid=np.zeros((ne, nd)), cface=np.zeros((ne, k), np.int64),
                 cadj=np.zeros((ne, k), np.int64), w=np.zeros((ne, k)), r=np.zeros((ne, k, nd)))
        _check(lib().bes_get_geometry(self.h, *[_p(o[x]) for x in
                                                ("owner", "nbr", "btype", "S", "n", "area", "vol", "centroid",
                                                 "cface", "cadj", "w", "r")]))
        return o

    # ---- state ----
    def set_batch(self, nbatch):
        """Advance nbatch freestream conditions (a power of two <= 64) together.

        Per-condition arrays (state, dt, gradient, limiter, residual, monitors)
        then carry a leading axis of length nbatch; the device state is discarded.
        """
        _check(lib().bes_set_batch(self.h, int(nbatch)))
        self.batch = int(nbatch)

    def set_limiter_freeze(self, drop):
        """Fix each condition's limiter once its density residual drop reaches
        `drop` in the implicit loop (0: never)."""
        _check(lib().bes_set_limiter_freeze(self.h, float(drop)))

    def freeze_limiter_values(self, phi):
        """Fix every condition's limiter at phi (the layout gradient_limiter returns)."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        assert phi.shape == self._shape(self.ncell, self.nv)
        _check(lib().bes_freeze_limiter_values(self.h, _p(phi)))

    def limiter_frozen(self):
        f = np.zeros(self.batch, np.int32)
        _check(lib().bes_limiter_frozen(self.h, _p(f)))
        return f.astype(bool)

    def set_freestream(self, primitive):
        """Freestream primitive [nv+1] for every condition, or [batch][nv+1]."""
        p = np.ascontiguousarray(primitive, np.float64)
        if p.ndim == 2:
            assert p.shape == (self.batch, self.nv + 1)
            _check(lib().bes_set_freestream_batch(self.h, _p(p)))
            return
        assert p.shape == (self.nv + 1,)
        _check(lib().bes_set_freestream(self.h, _p(p)))

    def set_state(self, U):
        U = np.ascontiguousarray(U, np.float64)
        assert U.shape == self._shape(self.ncell, self.nv)
        _check(lib().bes_set_state(self.h, _p(U)))

    def get_state(self):
        U = np.zeros(self._shape(self.ncell, self.nv))
        _check(lib().bes_get_state(self.h, _p(U)))
        return U

    # ---- passes ----
    def time_step(self, cfl, local=True):
        _check(lib().bes_compute_time_step(self.h, cfl, int(local)))
        dt = np.zeros(self._shape(self.ncell))
        _check(lib().bes_get_time_step(self.h, _p(dt)))
        return dt

    def gradient_limiter(self):
        _check(lib().bes_compute_gradient_limiter(self.h))
        G = np.zeros(self._shape(self.ncell, self.nv, self.ndim))
        P = np.zeros(self._shape(self.ncell, self.nv))
        _check(lib().bes_get_gradient(self.h, _p(G)))
        _check(lib().bes_get_limiter(self.h, _p(P)))
        return G, P

    def residual(self):
        """Second-order residual of the current state (gradient/limiter recomputed)."""
        _check(lib().bes_compute_gradient_limiter(self.h))
        R = np.zeros(self._shape(self.ncell, self.nv))
        mx = np.zeros(self._shape(self.nv))
        _check(lib().bes_compute_residual(self.h, _p(R), _p(mx)), self.h)
        return R, mx

    def write_solution(self, path, fmt="vtk", cond=0):
        """Solution file in the reference's legacy-VTK or CSV format."""
        _check(lib().bes_write_solution(self.h, str(path).encode(), 0 if fmt == "vtk" else 1, cond))

    def write_slip_wall(self, path, cond=0):
        """Slip-wall CSV in the reference's format."""
        _check(lib().bes_write_slip_wall(self.h, str(path).encode(), cond))

    def set_profiling(self, on=True):
        _check(lib().bes_set_profiling(self.h, int(on)))

    def autotune(self, explicit=False, implicit=False, coloring=False):
        """Block size per kernel kind (and the ILU(0) colouring) by timing on the
        current data; returns {kind: (block size, microseconds)} and the colouring."""
        flags = (1 if explicit else 0) | (2 if implicit else 0) | (4 if coloring else 0)
        ch = np.zeros(16, np.int32)
        us = np.zeros(16)
        col = C.c_int(-1)
        _check(lib().bes_autotune(self.h, flags, _p(ch), _p(us), C.byref(col)), self.h)
        out = {PROF_KINDS[k]: (int(ch[k]), float(us[k])) for k in range(len(PROF_KINDS)) if us[k] > 0}
        names = {v: k for k, v in self.COLORING.items()}
        return out, names.get(col.value)

    def profile(self):
        n = len(PROF_KINDS)
        ms = np.zeros(n)
        cnt = np.zeros(n, np.int64)
        by = np.zeros(n)
        _check(lib().bes_get_profile(self.h, _p(ms), _p(cnt), _p(by), n))
        return {k: dict(ms=float(ms[i]), launches=int(cnt[i]), bytes=float(by[i]),
                        GBps=float(by[i]) / (ms[i] * 1e-3) / 1e9 if ms[i] > 0 else 0.0)
                for i, k in enumerate(PROF_KINDS) if cnt[i] > 0}

    def synchronize(self):
        _check(lib().bes_synchronize(self.h))

    # ---- implicit path ----
    PRECOND = {"none": 0, "jacobi": 1, "ilu0": 2, "learned": 3, "amg": 4}
    AMG_RELAX = {"jacobi": 0, "spai0": 1, "block_jacobi": 2}

    def implicit_setup(self, precision=64, precond="ilu0", restart=30, maxiter=1000, rtol=1e-7):
        _check(lib().bes_implicit_setup(self.h, precision, self.PRECOND[precond], restart, maxiter, rtol))

    def implicit_set_reorth(self, eta):
        _check(lib().bes_implicit_set_reorth(self.h, eta))

    def implicit_stats(self):
        r = C.c_long()
        f = C.c_int()
        e = C.c_long()
        _check(lib().bes_implicit_stats(self.h, C.byref(r), C.byref(f), C.byref(e)))
        return {"reorths": r.value, "lp_fallbacks": f.value, "extra_restarts": e.value}

    SOLVER = {"gmres": 0, "fgmres": 0, "bicgstab": 1}
    COLORING = {"morton": 0, "mesh": 1, "flow": 2}

    def implicit_set_coloring(self, mode):
        """Greedy visiting order of the ILU(0) colouring: morton, mesh or flow."""
        _check(lib().bes_implicit_set_coloring(self.h, self.COLORING[mode]))

    def implicit_ncolors(self):
        return lib().bes_implicit_ncolors(self.h)

    def implicit_set_solver(self, solver):
        """Krylov method by the reference's LINEAR_SOLVER names (case-insensitive);
        FGMRES runs GMRES (the same iterates for a fixed preconditioner); LGMRES
        is rejected by the library."""
        code = self.SOLVER.get(str(solver).lower(), -1)
        _check(lib().bes_implicit_set_solver(self.h, code))

    def implicit_set_precond(self, precond):
        _check(lib().bes_implicit_set_precond(self.h, self.PRECOND[precond]))

    def implicit_assemble(self, cfl, local=True, repeat_nv=True):
        """Linf of b and of sum F*A per variable ([batch][nv] for a batch)."""
        lb = np.zeros(self._shape(self.nv))
        lr = np.zeros(self._shape(self.nv))
        _check(lib().bes_implicit_assemble(self.h, cfl, int(local), int(repeat_nv), _p(lb), _p(lr)), self.h)
        return lb, lr

    def implicit_system(self):
        """Assembled system in the reference CoefficientMatrix CSR layout."""
        nnz = lib().bes_implicit_nnz(self.h)
        rp = np.zeros(self.ncell * self.nv + 1, np.int64)
        col = np.zeros(nnz, np.int64)
        val = np.zeros(nnz)
        rhs = np.zeros((self.ncell, self.nv))
        _check(lib().bes_implicit_get_system(self.h, _p(rp), _p(col), _p(val), _p(rhs)))
        return rp, col, val, rhs

    def linear_solve(self, allow_unconverged=False):
        it = C.c_int()
        rr = C.c_double()
        st = lib().bes_linear_solve(self.h, C.byref(it), C.byref(rr))
        if not (allow_unconverged and st == ERR_SOLVE):
            _check(st, self.h)
        return it.value, rr.value

    def linear_solution(self):
        x = np.zeros(self._shape(self.ncell, self.nv))
        _check(lib().bes_get_linear_solution(self.h, _p(x)))
        return x

    def implicit_update(self):
        _check(lib().bes_implicit_update(self.h))

    def implicit_run(self, nsteps, cfl_min, cfl_max=0.0, exponents=(1.0, 1.0), local=True, stop_drop=0.0,
                     repeat_nv=True):
        """Backward-Euler steps; repeat_nv=True is the reference's nV-fold block repeat."""
        ex = np.ascontiguousarray(exponents, np.float64)
        hist = np.zeros(self._hist_shape(nsteps))
        done = C.c_int()
        _check(lib().bes_implicit_run(self.h, nsteps, cfl_min, cfl_max, _p(ex), int(local), int(repeat_nv),
                                      stop_drop, _p(hist), C.byref(done)), self.h)
        return hist[:done.value]

    def _hist_shape(self, nsteps):
        """Rows of the implicit loop's history: [nsteps][2nv+3], or [nsteps][batch][2nv+3]."""
        w = 2 * self.nv + 3
        return (nsteps, w) if self.batch == 1 else (nsteps, self.batch, w)

    def implicit_batch_state(self):
        """Per condition: steps taken, drop reached, next Courant number, and the
        Krylov iterations and true relative residual of the last solve."""
        B = self.batch
        steps = np.zeros(B, np.int32)
        reached = np.zeros(B, np.int32)
        cfl = np.zeros(B)
        iters = np.zeros(B, np.int32)
        rr = np.zeros(B)
        _check(lib().bes_implicit_batch_state(self.h, _p(steps), _p(reached), _p(cfl), _p(iters), _p(rr)))
        return {"steps": steps, "reached": reached.astype(bool), "cfl": cfl, "iters": iters, "relres": rr}

    # ---- learned preconditioner ----
    def lp_init(self, seed=20261202):
        _check(lib().bes_lp_init(self.h, seed))

    def lp_params(self):
        p = np.zeros(lib().bes_lp_param_count(), np.float32)
        _check(lib().bes_lp_get_params(self.h, _p(p), p.size))
        return p

    def lp_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().bes_lp_set_params(self.h, _p(p), p.size))

    def lp_operator(self):
        """Blocks of the learned operator M for the assembled system (reference order)."""
        nv, k = self.nv, self.k
        Md = np.zeros((self.ncell, nv, nv))
        Me = np.zeros((self.ncell, k, nv, nv))
        G = np.zeros((self.ncell, k, nv))
        _check(lib().bes_lp_operator(self.h, _p(Md), _p(Me), _p(G)), self.h)
        return Md, Me, G

    def lp_loss_grad(self, probes, nrich=3):
        V = np.ascontiguousarray(probes, np.float64)
        assert V.ndim == 3 and V.shape[1:] == (self.ncell, self.nv)
        loss = C.c_double()
        g = np.zeros(lib().bes_lp_param_count(), np.float32)
        _check(lib().bes_lp_loss_grad(self.h, _p(V), V.shape[0], nrich, C.byref(loss), _p(g)), self.h)
        return loss.value, g

    def lp_train(self, steps, lr=3e-3, nprobe=4, nrich=3, seed=1):
        hist = np.zeros(steps)
        _check(lib().bes_lp_train(self.h, steps, lr, nprobe, nrich, seed, _p(hist)), self.h)
        return hist

    def lp_set_online(self, steps, lr=3e-3, nprobe=4, nrich=3):
        """Adam steps on each assembled system before its solve (0: off)."""
        _check(lib().bes_lp_set_online(self.h, int(steps), float(lr), int(nprobe), int(nrich)))

    def lp_online_loss(self):
        return float(lib().bes_lp_online_loss(self.h))

    def lp_neutral(self):
        """Decoder bias 0: every blend starts at 1/2, no hand-set prior."""
        p = self.lp_params()
        p[-32:] = 0.0
        self.lp_set_params(p)

    def lp_fallbacks(self):
        return lib().bes_lp_fallbacks(self.h)

    def lp_set_rebuild_interval(self, n):
        _check(lib().bes_lp_set_rebuild_interval(self.h, int(n)))

    def lp_builds(self):
        f = lib().bes_lp_builds
        f.restype = C.c_long
        f.argtypes = [C.c_void_p]
        return f(self.h)

    # ---- aggregation AMG preconditioner ----
    def amg_set_params(self, eps_strong=0.08, over_interp=1.5, damping=0.72, coarse_enough=500, relax="jacobi",
                       rebuild_interval=1):
        _check(lib().bes_amg_set_params(self.h, eps_strong, over_interp, damping, coarse_enough,
                                        self.AMG_RELAX[relax], rebuild_interval))

    def amg_set_aggregation(self, mode):
        """"parallel" (default) or "row_sweep" (AMGCL's own serial sweep, for checks)."""
        _check(lib().bes_amg_set_aggregation(self.h, {"parallel": 0, "row_sweep": 1}[mode]))

    def amg_info(self, with_builds=False):
        """(rows per level, stored blocks per level, coarsest inverted[, builds so far])."""
        nl = C.c_int(0)
        direct = C.c_int(0)
        builds = C.c_long(0)
        rows = np.zeros(32, np.int64)
        nnzb = np.zeros(32, np.int64)
        _check(lib().bes_amg_info(self.h, C.byref(nl), _p(rows), _p(nnzb), C.byref(direct), C.byref(builds), 32))
        out = (rows[:nl.value].copy(), nnzb[:nl.value].copy(), bool(direct.value))
        return out + (builds.value,) if with_builds else out

    def amg_level(self, level):
        """Aggregates and states of the rows of a level, and for level >= 1 its
        block CSR (rowptr, col, val [nnzb][nv][nv]) or None for level 0."""
        rows, nnzb, _ = self.amg_info()
        n = int(rows[level])
        agg = np.zeros(n, np.int64)
        st = np.zeros(n, np.uint8)
        if level == 0:
            _check(lib().bes_amg_level(self.h, 0, _p(agg), _p(st), None, None, None))
            return agg, st, None
        rp = np.zeros(n + 1, np.int64)
        col = np.zeros(int(nnzb[level]), np.int64)
        val = np.zeros((int(nnzb[level]), self.nv, self.nv))
        _check(lib().bes_amg_level(self.h, level, _p(agg), _p(st), _p(rp), _p(col), _p(val)))
        return agg, st, (rp, col, val)

    def amg_apply(self, r):
        r = np.ascontiguousarray(r, np.float64).reshape(self.ncell, self.nv)
        z = np.zeros_like(r)
        _check(lib().bes_amg_apply(self.h, _p(r), _p(z)))
        return z

    def implicit_continue(self, nsteps):
        """nsteps more steps of the last implicit_run; returns (history, reached)."""
        hist = np.zeros(self._hist_shape(nsteps))
        done = C.c_int()
        reached = C.c_int()
        _check(lib().bes_implicit_continue(self.h, nsteps, _p(hist), C.byref(done), C.byref(reached)), self.h)
        return hist[:done.value], bool(reached.value)

    def explicit_graph(self, on=True):
        """Replay each pseudo-time step as one captured HIP graph."""
        _check(lib().bes_explicit_graph(self.h, int(on)))

    def explicit_persistent(self, on=True):
        """Run all steps of explicit_run in one cooperative persistent launch
        (one condition, local time steps)."""
        _check(lib().bes_explicit_persistent(self.h, int(on)))

    def explicit_run(self, niter, cfl, local=True, alpha=RK4_ALPHA, monitor=True):
        a = np.ascontiguousarray(alpha, np.float64)
        linf = np.zeros((niter,) + self._shape(self.nv)) if monitor else None
        _check(lib().bes_explicit_run(self.h, niter, len(a), _p(a), cfl, int(local), _p(linf)), self.h)
        return linf


def flux_batch(ndim, kind, UL, UR, normals, gas=None):
    gas = gas or default_gas()
    nv = ndim + 2
    UL = np.ascontiguousarray(UL, np.float64).reshape(-1, nv)
    n = UL.shape[0]
    URp = np.zeros((n, nv + 1))
    if UR is not None:
        UR = np.asarray(UR, np.float64).reshape(n, -1)
        URp[:, :UR.shape[1]] = UR
    nrm = np.ascontiguousarray(normals, np.float64).reshape(n, ndim)
    F = np.zeros((n, nv))
    _check(lib().bes_flux_batch(ndim, kind, n, _p(UL), _p(URp), _p(nrm), C.byref(gas), _p(F)))
    return F


def jacobian_batch(ndim, kind, UL, UR, normals, gas=None):
    gas = gas or default_gas()
    nv = ndim + 2
    UL = np.ascontiguousarray(UL, np.float64).reshape(-1, nv)
    n = UL.shape[0]
    URp = np.zeros((n, nv + 1))
    if UR is not None:
        UR = np.asarray(UR, np.float64).reshape(n, -1)
        URp[:, :UR.shape[1]] = UR
    nrm = np.ascontiguousarray(normals, np.float64).reshape(n, ndim)
    JL = np.zeros((n, nv, nv))
    JR = np.zeros((n, nv, nv))
    _check(lib().bes_jacobian_batch(ndim, kind, n, _p(UL), _p(URp), _p(nrm), C.byref(gas), _p(JL), _p(JR)))
    return JL, JR
