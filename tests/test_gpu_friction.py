"""The hybrid wall friction of the 3-D serial program (mps_fluid3D/src/
operators.c:337-463) in libmps_hip3ds.so, the library built in that
program's SimConfig layout.

- apply_friction, the reference's own entry point, against the fixture the
  reference wrote (golden/friction3ds.npz): bit for bit.  The contact angle
  goes through acos, whose last bit may differ between the host's libm and
  the device's; it only decides the static-friction test, at exact
  thresholds.
- Whole steps with the friction on against the oracle's serial build, to
  the tolerances of the other step tests (the CG's dot products sum in
  another order than the reference's).
"""
import ctypes as C
import sys

import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_apply_friction_matches_reference_fixture(lib3s):
    sc, cfg, want, fl = H.friction_fixture(lib3s.abi)
    lib3s.set_config(cfg)
    n = sc.num
    re = max(cfg.influence_radius_lap, cfg.influence_radius_n)
    dmin = (C.c_double * 3)(*[cfg.domain_min[d] for d in range(3)])
    dmax = (C.c_double * 3)(*[cfg.domain_max[d] for d in range(3)])
    nl = lib3s.neighbor_list_create(n, cfg.max_neighbors)
    cl = lib3s.cell_list_create(n, re, dmin, dmax)
    say("apply_friction on", n, "particles")
    lib3s.neighbor_search_cell_linked_list(nl, sc.ptr(), cl, re)
    before = np.array(sc.arr["vel"])[fl]
    lib3s.apply_friction(sc.ptr(), nl)
    assert lib3s.mps_hip_last_error() == 0
    got = np.array(sc.arr["vel"])[fl]
    lib3s.mps_hip_release(sc.ptr())
    lib3s.neighbor_list_free(nl)
    lib3s.cell_list_free(cl)
    bad = np.any(got.view(np.uint64) != want.view(np.uint64), axis=1)
    say("changed by the friction:", int(np.any(got != before, axis=1).sum()), "differing from the fixture:",
        int(bad.sum()))
    assert not bad.any()


@pytest.mark.parametrize("v_c", [0.01, 0.002])
def test_friction_steps_vs_oracle(lib3s, v_c):
    params = dict(friction_enabled=1, friction_delta_theta=30.0, dynamic_friction_alpha=600.0,
                  wetting_angle_SL=60.0, friction_velocity_threshold=v_c)
    dim, cfg, pos, vel, typ = H.build("dam3d_small", friction=True, cg_tolerance=1e-12, **params)
    lib3s.set_option("tol_mode", 1)
    say("friction steps, v_c", v_c, "N=%d" % len(typ))
    o = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1, friction=True)
    ob = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1, dot_order=1, friction=True)
    g = H.GpuRun(lib3s, cfg, pos, vel, typ, n0=o.scene.ps.n0, lam=o.scene.ps.lambda_)
    for s in range(1, 9):
        assert g.step() == 0
        o.step()
        ob.step()
        a, b, c = g.scene.arr, o.scene.arr, ob.scene.arr
        rec = dict(step=s, p_gpu=H.rel_err(a["pressure"], b["pressure"]),
                   p_spread=H.rel_err(c["pressure"], b["pressure"]),
                   x_gpu=H.rel_err(a["pos"], b["pos"]), x_spread=H.rel_err(c["pos"], b["pos"]),
                   type_exact=bool(np.array_equal(a["type"], b["type"])))
        say(rec)
        assert rec["type_exact"]
        assert rec["p_gpu"] <= max(100 * rec["p_spread"], 1e-9)
        assert rec["x_gpu"] <= max(100 * rec["x_spread"], 1e-13)
    g.close()
