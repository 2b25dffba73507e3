"""Writes small frozen outputs of the CPU oracle (plain-C FP64 restatement
of the reference's functors) as regression fixtures:
tests/golden/oracle_fixtures.npz.

Inputs: icosahedral (seed 0) and cubed-sphere (seed 1) meshes of level 2
from this package's generator, Rossby-Haurwitz 5-4 vorticity
(u0 = 0.1, amp = -0.1), Omega = 2 pi.  Outputs per mesh: vertex and face
velocity (faces collocated, divided faces skipped as sources), vertex
stream function, and the state after one RK4 step of dt = 0.01 (classical
RK4, unprojected stages).  Run:  python tests/golden/make_oracle_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_fixtures.npz")
LEVEL = 2
DT = 0.01
OMEGA = 2 * np.pi


def fixture_arrays(ora, seed):
    from lpm_amd import galleries as G
    from lpm_amd.mesh import SphereMesh
    m = SphereMesh.build(seed, LEVEL)
    vx = np.ascontiguousarray(m.vert_x.T)
    fx = np.ascontiguousarray(m.face_x.T)
    z = G.rh54_vorticity(m.face_x, 0.1, -0.1)
    zv = G.rh54_vorticity(m.vert_x, 0.1, -0.1)
    uv = ora.bve_velocity(vx, fx, z, m.face_area, m.face_mask)
    uf = ora.bve_velocity(fx, fx, z, m.face_area, m.face_mask, collocated=True)
    psi = ora.bve_stream_fn(vx, fx, z, m.face_area, m.face_mask)
    st = {"vx": vx.copy(), "fx": fx.copy(), "vzeta": zv.copy(), "fzeta": z.copy(),
          "farea": m.face_area.copy(), "fmask": m.face_mask.copy(),
          "vu": uv.copy(), "fu": uf.copy()}
    ora.bve_rk4_step(st, DT, OMEGA, compat=False, renorm=False)
    return {"vert_u": uv, "face_u": uf, "vert_psi": psi,
            "step_vert_x": st["vx"], "step_face_x": st["fx"],
            "step_vert_zeta": st["vzeta"], "step_face_zeta": st["fzeta"],
            "step_vert_u": st["vu"]}


def main():
    import oracle as ora
    ora.lib()
    arrays = {}
    for seed, tag in ((0, "icos"), (1, "cube")):
        for k, v in fixture_arrays(ora, seed).items():
            arrays[f"{tag}_{k}"] = np.asarray(v, dtype=np.float64)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}: {sorted(arrays)}")


if __name__ == "__main__":
    main()
