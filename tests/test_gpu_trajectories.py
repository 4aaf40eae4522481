"""Multi-step GPU trajectories against the CPU oracle and against the
reference's own shipped snapshots.

The CG's dot products sum in a different order on the GPU than in the
reference's serial loop (mps_fluid3D_parallel/src/pressure_solver.c:70-84),
so trajectories agree to a spread, not bit for bit.  The spread is measured,
not assumed: two oracle runs that differ only in their dot-product order
(forward and backward serial sums) show how far the reference itself moves
under a reordering.  At every step the GPU's relative error against the
oracle must stay within 100x that spread (or a floor of 1e-13 for
positions, 1e-9 for pressure and velocity, where the spread is below the
solver's own resolution).

Measured oracle spreads (relative L2, CG to 1e-12 relative):
  plain 2-D dam break, 2,395 particles: step 100 positions 2.5e-16,
    pressure 7.1e-14, velocities 1.8e-14
  3-D dam3d_small, 25,785 particles: step 100 positions 4.0e-17,
    pressure 6.7e-14, velocities 8.9e-15
Over the plain 2-D dam break's 1,000 steps (CG to 1e-9 relative) the
spread grows from 1e-16 to 3e-2 in positions once the column hits the far
wall (about step 500): the flow is chaotic there.  Until the position
spread reaches 1e-9 (about step 300) every step meets the full rule above
(types, pressure, positions, velocities).  Past it the two oracle runs and
the GPU flip different surface flags and lose different splashing
particles, so those steps compare positions and velocities to 100x the
spread, and the end state compares the fluid particle count and the fluid
centroid.  Measured on the GPU: position error 1.2-1.6x the
spread at steps 100-400, 14x at step 500 (impact), 1.2x at step 1000.
"""
import os

import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.gpu


def _check_records(recs, x_floor=1e-13, pv_floor=1e-9):
    for r in recs:
        assert r["type_exact"], r
        assert r["p_gpu"] <= max(100 * r["p_spread"], pv_floor), r
        assert r["x_gpu"] <= max(100 * r["x_spread"], x_floor), r
        assert r["v_gpu"] <= max(100 * r["v_spread"], pv_floor), r


@pytest.mark.parametrize("name,libname", [("dam2d_l0.01", "lib2"), ("dam3d_small", "lib3")])
def test_100_steps_vs_oracle(request, name, libname):
    """100 steps from rest (mps_fluid3D_parallel/src/simulation.c:19-118
    each), GPU against the oracle at every step."""
    lib = request.getfixturevalue(libname)
    g, o, recs = H.step_compare(lib, name, 100)
    _check_records(recs)
    last = recs[-1]
    H.say("step 100:", last)
    g.close()


def _centroid(arr):
    fl = arr["type"] == 0
    return arr["pos"][fl].mean(axis=0), int(fl.sum())


def test_config0_1000_steps(lib2):
    """The plain 2-D dam break of the benchmark's first configuration over
    its full 1,000 steps (dt 1e-3, CG to 1e-9 relative)."""
    dim, cfg, pos, vel, typ = H.build("dam2d_l0.01", cg_tolerance=1e-9)
    lib2.set_option("tol_mode", 1)
    o = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1)
    ob = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1, dot_order=1)
    g = H.GpuRun(lib2, cfg, pos, vel, typ, n0=o.scene.ps.n0, lam=o.scene.ps.lambda_)
    recs = []
    for s in range(1, 1001):
        assert g.step() == 0
        o.step()
        ob.step()
        recs.append(H.compare_states(s, g.scene.arr, o.scene.arr, ob.scene.arr))
        if s % 100 == 0:
            H.say(recs[-1])
    pre = [r for r in recs if r["x_spread"] < 1e-9]
    post = [r for r in recs if r["x_spread"] >= 1e-9]
    H.say("%d steps before the spread reaches 1e-9, %d after" % (len(pre), len(post)))
    assert len(pre) >= 250
    _check_records(pre)
    for r in post:
        assert r["x_gpu"] <= 100 * r["x_spread"], r
        assert r["v_gpu"] <= 100 * r["v_spread"], r
    worst = max(recs, key=lambda r: r["x_gpu"] / max(r["x_spread"], 1e-13))
    H.say("largest position error over the oracle's spread:", worst)
    cg, ng = _centroid(g.scene.arr)
    co, no = _centroid(o.scene.arr)
    cb, nb = _centroid(ob.scene.arr)
    H.say("fluid centroid gpu", cg, ng, "oracle", co, no, "oracle (other dot order)", cb, nb)
    # splashing particles leave the tank (ghosts) in one realisation and not
    # the other; the fluid count and centroid are compared at that scale
    assert abs(ng - no) <= max(2 * abs(nb - no), 2)
    assert np.linalg.norm(cg - co) <= max(10 * np.linalg.norm(cb - co), 1e-6)
    g.close()


def _printed(a):
    """Values as the reference's "%.8e" output reads back."""
    return np.array([float("%.8e" % v) for v in np.ravel(a)]).reshape(np.shape(a))


def test_shipped_3d_dambreak_vs_reference_snapshots(lib3):
    """The reference's shipped 3-D dam break (examples/dambreak: 28,112
    particles, ICCG, hs_mode 1, clamp, cg_tolerance 1e-8 absolute) run 200
    steps on the GPU, against the reference's own snapshots at steps 100
    and 200 (output/output_000100.csv, _000200.csv, stored in
    golden/shipped3d_output.npz).  The snapshots hold 9 significant digits
    ("%.8e", io.c:10-34).  The oracle run here matches them to that
    printing: step 100 with every position equal after rounding, step 200
    with one position 1e-11 apart, pressure within 7e-12 relative.
    The GPU's values are rounded the same way before comparing.  Bound:
    each fluid coordinate within 2e-8 relative (two units of the last
    printed digit), pressure within 1e-9 relative, types equal."""
    g = np.load(os.path.join(H.GOLDEN, "shipped3d_output.npz"))
    over = dict(solver_type=1, hs_mode=1, influence_ratio_lap=2.1, influence_ratio_n=2.1,
                wall_layers=2, dummy_layers=2, cg_tolerance=1e-8, cg_max_iter=10000,
                clamp_negative_pressure=1)
    dim, cfg, pos, vel, typ = H.build("shipped3d", **over)
    lib3.set_option("tol_mode", 0)
    try:
        run = H.GpuRun(lib3, cfg, pos, vel, typ)
        fi = g["fluid_index"]
        for s in range(1, 201):
            assert run.step() == 0
            if s in (100, 200):
                a = run.scene.arr
                ref_pos = g["step%d_fluid_pos" % s]
                pos8 = _printed(a["pos"][fi])
                dx = np.abs(pos8 - ref_pos)
                rel = dx / np.maximum(np.abs(ref_pos), 1e-30)
                perr = H.rel_err(_printed(a["pressure"]), g["step%d_pressure" % s])
                H.say("step", s, "coordinates differing after %%.8e: %d of %d, max rel %.3e, max abs %.3e, "
                      "pressure rel %.3e" % (int((dx > 0).sum()), dx.size, rel.max(), dx.max(), perr))
                assert np.array_equal(a["type"], g["step%d_type" % s].astype(np.int32))
                assert np.all(dx <= 2e-8 * np.abs(ref_pos) + 1e-15)
                assert perr <= 1e-9
        run.close()
    finally:
        lib3.set_option("tol_mode", 1)
