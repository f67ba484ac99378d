"""bes_flux_batch_noalloc (no allocation, fixed per-GPU staging) against
bes_flux_batch through the C ABI: the same face fluxes, bit for bit, on two
calls in a row (the second smaller), for every face kind in 2D and 3D, and
the size limit reported as BES_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

from bes_amd import core

pytestmark = pytest.mark.gpu

MAX_FACES = 16384


def _lib():
    L = core.lib()
    for name in ("bes_flux_batch", "bes_flux_batch_noalloc"):
        fn = getattr(L, name)
        fn.argtypes = [C.c_int, C.c_int, C.c_long, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        fn.restype = C.c_int
    return L


def _inputs(nd, kind, n, seed):
    rng = np.random.default_rng(seed)
    nv = nd + 2
    rho = rng.uniform(0.5, 1.5, n)
    vel = rng.uniform(-0.9, 0.9, (n, nd))
    p = rng.uniform(0.4, 1.2, n)
    UL = np.empty((n, nv))
    UL[:, 0] = rho
    UL[:, 1:1 + nd] = rho[:, None] * vel
    UL[:, nv - 1] = p / 0.4 + 0.5 * rho * (vel ** 2).sum(1)
    UR = np.zeros((n, nv + 1))
    if kind == core.INTERNAL:
        UR[:, :nv] = UL[rng.permutation(n)]
    elif kind == core.FREESTREAM:
        UR[:] = core.freestream_primitive(rng.uniform(-0.9, 0.9, nd), nd)
    nrm = rng.normal(size=(n, nd))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return [np.ascontiguousarray(a) for a in (UL, UR, nrm)]


def _call(L, fn, nd, kind, UL, UR, nrm):
    n = UL.shape[0]
    F = np.full((n, nd + 2), np.nan)
    g = core.default_gas()
    st = getattr(L, fn)(nd, kind, n, UL.ctypes.data, UR.ctypes.data if kind != core.SLIP_WALL else None,
                        nrm.ctypes.data, C.addressof(g), F.ctypes.data)
    return st, F


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("kind", [core.INTERNAL, core.SLIP_WALL, core.FREESTREAM])
def test_noalloc_flux_equals_flux_batch(nd, kind):
    L = _lib()
    print("case", nd, kind, flush=True)
    for call, n in enumerate((MAX_FACES, 3001)):   # second call smaller than the first
        UL, UR, nrm = _inputs(nd, kind, n, seed=100 * nd + 10 * kind + call)
        st0, F0 = _call(L, "bes_flux_batch", nd, kind, UL, UR, nrm)
        st1, F1 = _call(L, "bes_flux_batch_noalloc", nd, kind, UL, UR, nrm)
        assert st0 == core.OK and st1 == core.OK
        assert np.isfinite(F0).all()
        assert np.array_equal(F0, F1), f"call {call}: noalloc flux differs"


def test_noalloc_rejects_oversized_batch():
    L = _lib()
    UL, UR, nrm = _inputs(3, core.INTERNAL, MAX_FACES + 1, seed=7)
    st, _ = _call(L, "bes_flux_batch_noalloc", 3, core.INTERNAL, UL, UR, nrm)
    assert st == core.ERR_ARG
    assert b"exceeds" in L.bes_last_error_message()
