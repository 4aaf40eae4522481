"""Writes tests/golden/friction3ds.npz, the fixture of the hybrid wall
friction of the 3-D serial program (mps_fluid3D/src/operators.c:337-463).

The small tank `fric3d_tiny` runs 8 steps in the oracle's serial layout
with the friction on; then the reference's own apply_friction (mps_fluid3D
built from its sources into oracle/_ref/libref3d.so by oracle/Makefile) is
applied to that state, over the reference's own neighbour list.  The
oracle's apply_friction must give the same bits here.  Stored: the fluid
particles' positions and velocities before, their velocities after, the
types, n0/lambda and the parameters; wall positions are the generator's.

Run from the repository root where the reference build exists:
    python tests/golden/make_friction_fixture.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import harness as H  # noqa: E402
from mps_amd import abi as A  # noqa: E402
from oracle.oracle import Oracle, Reference  # noqa: E402

SCENE = "fric3d_tiny"
STEPS = 8
PARAMS = dict(friction_enabled=1, friction_delta_theta=30.0, dynamic_friction_alpha=600.0,
              wetting_angle_SL=60.0, friction_velocity_threshold=0.002, cg_tolerance=1e-10)


def main():
    dim, cfg, pos, vel, typ = H.build(SCENE, friction=True, **PARAMS)
    run = H.OracleRun(cfg, pos, vel, typ, dim, friction=True)
    for _ in range(STEPS):
        run.step()
    state = run.scene.arr.copy()
    n0, lam = run.scene.ps.n0, run.scene.ps.lambda_
    # the reference's apply_friction on that state
    ref = Reference("3d")
    cfg_r = ref.config_from(cfg)
    ref.set_config(cfg_r)
    sr = A.HostScene(ref.abi, state["pos"], state["vel"], state["type"])
    sr.arr[:] = state
    sr.ps.n0, sr.ps.lambda_ = n0, lam
    nl, cl = ref.make_lists(sr.num)
    ref.L.neighbor_search_cell_linked_list(nl, sr.ptr(), cl, ref.search_radius())
    ref.L.apply_friction.argtypes = [C.POINTER(ref.abi.ParticleSystem), C.POINTER(ref.abi.NeighborList)]
    ref.L.apply_friction(sr.ptr(), nl)
    ref.free_lists(nl, cl)
    vout = np.array(sr.arr["vel"])
    # the oracle's, on the same state
    so = A.HostScene(A.types(3, True), state["pos"], state["vel"], state["type"])
    so.arr[:] = state
    so.ps.n0, so.ps.lambda_ = n0, lam
    Oracle(3, True).apply_friction(so, cfg)
    assert np.array_equal(np.array(so.arr["vel"]).view(np.uint64), vout.view(np.uint64))
    fl = state["type"] == 0
    changed = int(np.any(vout[fl] != state["vel"][fl], axis=1).sum())
    assert changed > 0
    print("%d particles, %d fluid, friction changed %d" % (len(state), fl.sum(), changed))
    np.savez_compressed(os.path.join(HERE, "friction3ds.npz"), scene=np.array(SCENE), steps=np.array(STEPS),
                        params=np.array(json.dumps(PARAMS)), type=np.array(state["type"], dtype=np.int8),
                        fluid_pos=np.array(state["pos"][fl]), fluid_vel_in=np.array(state["vel"][fl]),
                        fluid_vel_out=vout[fl], n0=np.array(n0), lam=np.array(lam))


if __name__ == "__main__":
    main()
