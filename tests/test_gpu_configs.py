"""The benchmark workloads themselves, on the GPU, against the oracle.

- 3-D dam break at l0 = 0.005 m (1,077,505 particles, 594,000 fluid):
  the GPU's initial parameters, its initial neighbour lists (set and
  order) and, after one step of CG to a relative residual of 1e-9, its
  lists, number density, flags, types and PPE (unknown numbering,
  pattern, values, right-hand side) bit for bit against the oracle;
  pressure, positions and velocities within 100x the spread between two
  oracle runs that differ only in their dot-product order.  The ICCG leg
  of the same scene: the IC(0) factor bit for bit.
- The same scene refined to l0 = 0.0035 m (2,667,690 particles): the
  mixed form (FP32 vectors, coded off-diagonals, FP64 refinement) against
  the FP64 GPU run, 3 steps, relative L2 pressure and position error at
  most 1e-6.
- The seeded 64-scene ensemble: the batched run equal bit for bit to
  single-scene runs of every scene, and scene 0 after one step against
  the oracle.

The oracle runs here are the largest of the suite (one CG step of the
1.08M-particle scene takes 15-30 s of host time per oracle run).
"""
import ctypes as C
import sys

import numpy as np
import pytest

import harness as H
from mps_amd import abi as A
from mps_amd import params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lists_match(nl, n, start, idx):
    """The padded NeighborList `nl` (n rows) holds exactly the rows of the
    oracle's CSR lists (start, idx), each in the same order."""
    cnt = np.ctypeslib.as_array(nl.count, shape=(n,))
    if not np.array_equal(cnt.astype(np.int64), np.diff(start)):
        return False
    mx = nl.max_neighbors
    flat = np.ctypeslib.as_array(nl.neighbors, shape=(n * mx,)).reshape(n, mx)
    mask = np.arange(mx, dtype=np.int32)[None, :] < cnt[:, None]
    return bool(np.array_equal(flat[mask], idx))


def test_config1_dam3d_l0005_first_step(lib3):
    name = "dam3d_l0.005"
    dim, cfg, pos, vel, typ = H.build(name, cg_tolerance=1e-9)
    n = len(typ)
    assert n == 1077505 and int((typ == 0).sum()) == 594000
    lib3.set_option("tol_mode", 1)
    say("config 1:", name, "N=%d" % n)
    g = H.GpuRun(lib3, cfg, pos, vel, typ)
    try:
        o = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1)
        for f in H.INIT_FIELDS:
            assert getattr(g.scene.ps, f) == getattr(o.scene.ps, f), f
        st, ix = o.sim.neighbor_lists()
        say("initial lists: %d entries" % len(ix))
        assert _lists_match(g.nl.contents, n, st, ix)
        ob = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1, dot_order=1)
        assert g.step() == 0
        o.step()
        ob.step()
        r = H.compare_states(1, g.scene.arr, o.scene.arr, ob.scene.arr)
        say(r, "iterations", lib3.stats(g.scene).last_iters)
        assert r["n_exact"] and r["flags_exact"] and r["type_exact"]
        st, ix = o.sim.neighbor_lists()
        assert _lists_match(g.nl.contents, n, st, ix)
        q = H.assert_ppe_exact(lib3, g, o)
        say("PPE: %d unknowns, %d entries" % (q["n_eq"], q["nnz"]))
        assert r["p_gpu"] <= max(100 * r["p_spread"], 1e-10)
        assert r["x_gpu"] <= max(100 * r["x_spread"], 1e-14)
        assert r["v_gpu"] <= max(100 * r["v_spread"], 1e-10)
    finally:
        g.close()


def test_config1_iccg_factor(lib3):
    """The ICCG leg of the 1.08M-particle scene: the PPE of step 1 (bit for
    bit the reference's, as the CG test shows) factored on the GPU equals
    the oracle's IC(0) factor bit for bit, and the two ICCG pressures
    agree to the solve's tolerance."""
    name = "dam3d_l0.005"
    dim, cfg, pos, vel, typ = H.build(name, cg_tolerance=1e-9, solver_type=1)
    lib3.set_option("tol_mode", 1)
    say("config 1 ICCG:", name, "N=%d" % len(typ))
    g = H.GpuRun(lib3, cfg, pos, vel, typ)
    try:
        o = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1)
        assert g.step() == 0
        o.step()
        rp, col, val, dinv = _export_ic(lib3, g.scene)
        oq = o.ppe()
        say("factor: %d rows, %d lower entries" % (len(dinv), len(col)))
        assert np.array_equal(rp, oq["l_ptr"].astype(np.int32))
        assert np.array_equal(col, oq["l_col"])
        assert np.array_equal(val.view(np.uint64), oq["l_val"].view(np.uint64))
        assert np.array_equal(dinv.view(np.uint64), oq["l_diag"].view(np.uint64))
        ep = H.rel_err(g.scene.arr["pressure"], o.scene.arr["pressure"])
        say("ICCG pressure rel err %.3e, iterations gpu %d" % (ep, lib3.stats(g.scene).last_iters))
        assert ep <= 1e-6
    finally:
        g.close()


def _export_ic(lib, scene):
    st = lib.stats(scene)
    ne = st.n_eq
    lnnz = (st.nnz - ne) // 2
    rp = np.zeros(ne + 1, np.int32)
    col = np.zeros(max(lnnz, 1), np.int32)
    val = np.zeros(max(lnnz, 1))
    dinv = np.zeros(max(ne, 1))
    P = lambda a, t=C.c_double: a.ctypes.data_as(C.POINTER(t))
    lib.check(lib.mps_hip_export_ic(scene.ptr(), P(rp, C.c_int), P(col, C.c_int), P(val), P(dinv)),
              "export_ic")
    return rp, col[:lnnz], val[:lnnz], dinv[:ne]


def _gen(lib, w):
    dims = (C.c_double * 6)(*w["dims"])
    n = lib.mps_gen_dambreak(dims, w["l0"], 3, 2, None, None)
    ps = lib.particle_system_create(n)
    assert lib.mps_gen_dambreak(dims, w["l0"], 3, 2, ps, None) == n
    return ps, n


def _arr(lib, ps, n):
    return np.ctypeslib.as_array(C.cast(ps.contents.particles, C.POINTER(lib.abi.Particle)), shape=(n,))


def test_config2_mixed_vs_fp64(lib3):
    """2,667,690 particles, 3 steps: the mixed form (precision 3: FP32 CG
    vectors, distance-coded FP32 off-diagonals, FP64 residual, dot
    products and right-hand side in the refinement) against FP64."""
    w = dict(dims=(1.0, 0.6, 0.5, 0.3, 0.5, 0.5), l0=0.0035, dt=5e-4)
    cfg = params.bench_config(lib3, w, 0, 1e-9)
    lib3.set_option("tol_mode", 1)
    lib3.set_config(cfg)
    runs = {}
    try:
        for prec in (0, 3):
            ps, n = _gen(lib3, w)
            runs[prec] = (ps, n)
        ps0, n = runs[0]
        assert n == 2667690
        assert int((_arr(lib3, ps0, n)["type"] == 0).sum()) == 1701870
        say("config 2: N=%d" % n)
        lib3.particle_system_calc_initial_params(ps0)
        psm = runs[3][0]
        for f in ("n0", "lambda_", "C_LL", "C_SL", "n0_count"):
            setattr(psm.contents, f, getattr(ps0.contents, f))
        for s in range(1, 4):
            out = {}
            for prec in (0, 3):
                ps, _ = runs[prec]
                lib3.set_option("precision", prec)
                lib3.check(lib3.mps_hip_run_steps(ps, 1), "step")
                lib3.check(lib3.mps_hip_download(ps, None), "download")
                a = _arr(lib3, ps, n)
                out[prec] = (np.array(a["pressure"]), np.array(a["pos"]), lib3.stats(ps))
            ep = H.rel_err(out[3][0], out[0][0])
            ex = H.rel_err(out[3][1], out[0][1])
            say("step %d: p err %.3e pos err %.3e iters fp64 %d mixed %d"
                % (s, ep, ex, out[0][2].last_iters, out[3][2].last_iters))
            assert ep <= 1e-6
            assert ex <= 1e-6
            st = out[3][2]
            assert st.last_residual <= 1e-9 * st.last_bnorm * (1 + 1e-12)
    finally:
        lib3.set_option("precision", 0)
        for ps, _ in runs.values():
            lib3.mps_hip_release(ps)
            lib3.particle_system_free(ps)


def test_config3_ensemble64(lib3):
    """The seeded 64-scene ensemble (tank 0.8 x 0.5 x 0.4 m, columns of
    uneven size, per-scene viscosity): 2 batched steps equal 2 steps of
    each scene alone, bit for bit, with the same CG iteration counts; and
    scene 0's first step against the oracle (n, flags and types exact,
    pressure and positions within 100x the oracle's dot-order spread)."""
    ws = params.ensemble(64)
    steps = 2
    lib3.set_option("tol_mode", 1)

    def make(w):
        cfg = params.bench_config(lib3, w, 0, 1e-9)
        ps, n = _gen(lib3, w)
        lib3.set_config(cfg)
        lib3.particle_system_calc_initial_params(ps)
        return cfg, ps, n

    def fields(ps, n):
        a = _arr(lib3, ps, n)
        return {f: np.ascontiguousarray(a[f]).tobytes() for f in H.FIELDS}

    single = []
    scene0 = None
    for q, w in enumerate(ws):
        cfg, ps, n = make(w)
        if q == 0:
            a = _arr(lib3, ps, n)
            scene0 = (cfg, np.array(a["pos"]), np.array(a["vel"]), np.array(a["type"]),
                      ps.contents.n0, ps.contents.lambda_)
        lib3.set_config(cfg)
        lib3.check(lib3.mps_hip_run_steps(ps, steps), "run_steps")
        lib3.check(lib3.mps_hip_download(ps, None), "download")
        single.append((fields(ps, n), lib3.stats(ps).total_iters))
        lib3.mps_hip_release(ps)
        lib3.particle_system_free(ps)
    made = [make(w) for w in ws]
    total = sum(m[2] for m in made)
    say("ensemble: %d scenes, %d particles (%d-%d)" % (len(made), total, min(m[2] for m in made),
                                                       max(m[2] for m in made)))
    PSP = C.POINTER(lib3.abi.ParticleSystem)
    ps_arr = (PSP * len(made))(*[m[1] for m in made])
    cf_arr = (C.POINTER(lib3.abi.SimConfig) * len(made))(*[C.pointer(m[0]) for m in made])
    try:
        lib3.check(lib3.mps_hip_run_batch(ps_arr, cf_arr, len(made), steps), "run_batch")
        for q, (cfg, ps, n) in enumerate(made):
            lib3.check(lib3.mps_hip_download(ps, None), "download")
            got = fields(ps, n)
            for f in H.FIELDS:
                assert got[f] == single[q][0][f], (q, f)
            assert lib3.stats(ps).total_iters == single[q][1], q
    finally:
        for cfg, ps, n in made:
            lib3.mps_hip_release(ps)
            lib3.particle_system_free(ps)
    # scene 0, one step, against the oracle
    cfg, pos, vel, typ, n0, lam = scene0
    say("scene 0 vs oracle: N=%d" % len(typ))
    g = H.GpuRun(lib3, cfg, pos, vel, typ, n0=n0, lam=lam)
    try:
        o = H.OracleRun(cfg, pos, vel, typ, 3, tol_mode=1, n0=n0, lam=lam)
        ob = H.OracleRun(cfg, pos, vel, typ, 3, tol_mode=1, dot_order=1, n0=n0, lam=lam)
        assert g.step() == 0
        o.step()
        ob.step()
        r = H.compare_states(1, g.scene.arr, o.scene.arr, ob.scene.arr)
        say(r)
        assert r["n_exact"] and r["flags_exact"] and r["type_exact"]
        H.assert_ppe_exact(lib3, g, o)
        assert r["p_gpu"] <= max(100 * r["p_spread"], 1e-10)
        assert r["x_gpu"] <= max(100 * r["x_spread"], 1e-14)
    finally:
        g.close()
