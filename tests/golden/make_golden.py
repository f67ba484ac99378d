"""Generate the golden fixtures under tests/golden/ from the reference itself.

Runs in the build container only (the reference tree is not on the GPU box):
the reference's Boost-free translation units are compiled by ``oracle/Makefile``
(target ``ref``) into ``oracle/_ref/bes_ref_driver``; this script writes each
case's mesh (MSH2), configuration (.besi), state and known-answer inputs,
runs the driver, and stores inputs and reference outputs together as one
compressed ``.npz`` per case.  The fixtures are data only.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from bes_amd import meshgen as mg  # noqa: E402
from oracle import oracle as orc  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "bes_ref_driver")
OUT = os.path.dirname(os.path.abspath(__file__))


def kat_inputs(ndim, seed):
    """Flux/Jacobian known-answer records [kind, UL(nv), UR(nv+1), n(nd)]."""
    rng = np.random.default_rng(seed)
    nv = ndim + 2
    recs = []

    def unit(v):
        return v / np.linalg.norm(v)

    def state(mach_vec, rho=1.0, p=1.0 / 1.4):
        P = np.zeros(ndim + 3)
        P[0] = rho
        P[1:1 + ndim] = mach_vec
        P[ndim + 1] = p
        P[ndim + 2] = p / (rho * orc.GAS_R)
        return orc.conserved_from_primitive(P, ndim)

    for i in range(48):   # Roe: generic, near-equal and Harten-active pairs
        n = unit(rng.normal(size=ndim))
        if i % 4 == 0:     # normal velocity ~ 0 -> |V| below Harten delta
            t = rng.normal(size=ndim); t -= t.dot(n) * n
            ul = state(0.3 * unit(t) + 0.01 * rng.normal() * n, rho=1.0 + 0.1 * rng.normal())
            ur = state(0.3 * unit(t) + 0.01 * rng.normal() * n, rho=1.0 + 0.1 * rng.normal())
        elif i % 4 == 1:   # normal velocity ~ sound speed -> |V - c| small
            ul = state((1.0 + 0.02 * rng.normal()) * n)
            ur = state((1.0 + 0.02 * rng.normal()) * n, rho=1.0 + 0.05 * rng.normal())
        elif i % 4 == 2:   # near-equal states
            base = state(rng.uniform(-0.9, 0.9, size=ndim))
            ul = base * (1 + 1e-3 * rng.normal(size=nv))
            ur = base * (1 + 1e-3 * rng.normal(size=nv))
        else:              # generic
            ul = state(rng.uniform(-1.5, 1.5, size=ndim), rho=rng.uniform(0.5, 2.0),
                       p=rng.uniform(0.4, 1.2))
            ur = state(rng.uniform(-1.5, 1.5, size=ndim), rho=rng.uniform(0.5, 2.0),
                       p=rng.uniform(0.4, 1.2))
        recs.append(np.concatenate([[1.0], ul, ur, [0.0], n]))
    for i in range(12):   # slip wall
        n = unit(rng.normal(size=ndim))
        ul = state(rng.uniform(-1, 1, size=ndim), rho=rng.uniform(0.5, 2), p=rng.uniform(0.4, 1.2))
        recs.append(np.concatenate([[2.0], ul, np.zeros(nv + 1), n]))
    for i in range(48):   # far field: all four characteristic branches
        n = unit(rng.normal(size=ndim))
        minf = rng.choice([0.3, 0.84, 1.5, 2.2]) * unit(rng.normal(size=ndim))
        pinf = orc.freestream_primitive(minf, ndim)
        vn = rng.choice([-2.0, -0.5, 0.5, 2.0]) + 0.1 * rng.normal()
        t = rng.normal(size=ndim); t -= t.dot(n) * n
        ul = state(vn * n + 0.3 * t, rho=rng.uniform(0.8, 1.2), p=rng.uniform(0.6, 0.8))
        recs.append(np.concatenate([[3.0], ul, pinf, n]))
    return np.array(recs)


CASES = {
    "tri2d": dict(mesh=lambda: mg.tri_patch(5, 4, seed=3), mach=(0.5, 0.1), iters=3, seed=11),
    "bump2d": dict(mesh=lambda: mg.bump_channel(16, 8, seed=5), mach=(0.5, 0.0), iters=3, seed=12),
    "box3d": dict(mesh=lambda: mg.box_tets(3, 3, 2, seed=7),
                  mach=(0.84 * np.cos(np.radians(3.06)), 0.0, 0.84 * np.sin(np.radians(3.06))),
                  iters=2, seed=13),
    "box3d_super": dict(mesh=lambda: mg.box_tets(3, 2, 2, seed=8, slip_faces=("ymin", "zmin")),
                        mach=(1.6, 0.2, 0.1), iters=2, seed=14),
}


def run_case(name, spec, tmp):
    md = spec["mesh"]()
    nd = md.ndim
    d = os.path.join(tmp, name)
    os.makedirs(d)
    mg.write_msh2(os.path.join(d, "m.msh"), md)
    mg.write_besi(os.path.join(d, "c.besi"), os.path.join(d, "m.msh"), nd, mach=spec["mach"],
                  cfl=0.8, local_dt=True)
    U = orc.perturbed_state(md.ncell, nd, spec["mach"], spec["seed"])
    U.tofile(os.path.join(d, "state.bin"))
    kin = kat_inputs(nd, spec["seed"] + 100)
    kin.tofile(os.path.join(d, "kat_in.bin"))
    with open(os.path.join(d, "explicit_iters.txt"), "w") as fh:
        fh.write(str(spec["iters"]))
    subprocess.run([DRIVER, "c.besi", "."], cwd=d, check=True, stdout=subprocess.DEVNULL)
    nv = nd + 2
    meta = np.fromfile(os.path.join(d, "meta.bin"), np.int64)
    _, _, ne, nf, k = [int(x) for x in meta]

    def rd(fn, dt=np.float64):
        return np.fromfile(os.path.join(d, fn), dt)

    rec = dict(
        ndim=nd, mach=np.asarray(spec["mach"], np.float64), cfl=0.8, niter=spec["iters"],
        nodes=md.nodes, cells=md.cells, bfaces=md.bfaces, btype=md.btype, state=U, kat_in=kin,
        face_owner=rd("face_owner.bin", np.int64), face_nbr=rd("face_nbr.bin", np.int64),
        face_btype=rd("face_btype.bin", np.int64), face_S=rd("face_S.bin").reshape(nf, nd),
        face_n=rd("face_n.bin").reshape(nf, nd), face_area=rd("face_area.bin"),
        face_center=rd("face_center.bin").reshape(nf, nd), cell_vol=rd("cell_vol.bin"),
        cell_centroid=rd("cell_centroid.bin").reshape(ne, nd),
        cell_faces=rd("cell_faces.bin", np.int64).reshape(ne, k),
        cell_adj=rd("cell_adj.bin", np.int64).reshape(ne, k), cell_w=rd("cell_w.bin").reshape(ne, k),
        cell_r=rd("cell_r.bin").reshape(ne, k, nd), cell_rsel=rd("cell_rsel.bin", np.int64).reshape(ne, k),
        grad=rd("grad.bin").reshape(ne, nv, nd), lim=rd("lim.bin").reshape(ne, nv),
        dt_local=rd("dt_local.bin"), dt_global=rd("dt_global.bin"),
        face_flux=rd("face_flux.bin").reshape(nf, nv), resid=rd("resid.bin").reshape(ne, nv),
        resid_scale=rd("resid_scale.bin").reshape(ne, nv),
        csr_rowptr=rd("csr_rowptr.bin", np.int64), csr_col=rd("csr_col.bin", np.int64),
        csr_val=rd("csr_val.bin"), impl_rhs=rd("impl_rhs.bin").reshape(ne, nv),
        explicit_state=rd("explicit_state.bin").reshape(ne, nv),
        explicit_linf=rd("explicit_linf.bin").reshape(-1, nv),
        kat_flux=rd("kat_flux.bin").reshape(-1, nv), kat_JL=rd("kat_JL.bin").reshape(-1, nv, nv),
        kat_JR=rd("kat_JR.bin").reshape(-1, nv, nv),
    )
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **rec)
    print(f"{name}: ncell={ne} nface={nf} kat={kin.shape[0]}")


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    with tempfile.TemporaryDirectory() as tmp:
        for name, spec in CASES.items():
            run_case(name, spec, tmp)


if __name__ == "__main__":
    main()
