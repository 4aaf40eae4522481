"""GPU edge cases against the CPU oracle: halos too large for LDS (32-bit
columns next to 12-bit halo indices in one matrix), neighbour-list
overflow reported as an error code, a NaN particle turned ghost, and a
scene whose PPE has no unknowns, alone and inside a batch.

The reference exits on a neighbour-list overflow
(neighbor_search.c:221-227); with option exit_on_error 0 the library
returns MPS_ERR_NEIGHBOR_OVERFLOW instead, naming the smallest original
index of a particle whose list overflowed -- the first one the reference's
serial loop over i reaches.  Every write stays inside the list: a
particle's entries past max_neighbors are counted, not stored.  (The
lists themselves are not handed back after an overflow; the reference has
exited by then.)
"""
import ctypes as C
import re
import sys

import numpy as np
import pytest

import harness as H
from mps_amd import abi as A

pytestmark = pytest.mark.gpu

MPS_ERR_NEIGHBOR_OVERFLOW = 4
GHOST = 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare_steps(lib, cfg, dim, pos, vel, typ, steps):
    lib.set_option("tol_mode", 1)
    o = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1)
    ob = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1, dot_order=1)
    g = H.GpuRun(lib, cfg, pos, vel, typ, n0=o.scene.ps.n0, lam=o.scene.ps.lambda_)
    recs = []
    for s in range(1, steps + 1):
        assert g.step() == 0
        o.step()
        ob.step()
        a, b, c = g.scene.arr, o.scene.arr, ob.scene.arr
        rec = dict(step=s,
                   p_gpu=H.rel_err(a["pressure"], b["pressure"]),
                   p_spread=H.rel_err(c["pressure"], b["pressure"]),
                   x_gpu=H.rel_err(a["pos"], b["pos"]),
                   x_spread=H.rel_err(c["pos"], b["pos"]),
                   n_exact=bool(np.array_equal(a["n"].view(np.uint64), b["n"].view(np.uint64))),
                   type_exact=bool(np.array_equal(a["type"], b["type"])))
        say(rec)
        recs.append(rec)
    return g, o, recs


def test_wide_halos_global_columns(lib3):
    """influence_ratio_lap 8: a 256-row block of the small dam break then
    references up to ~7,000 unknowns, more than an LDS halo holds (4,096),
    so about a quarter of the blocks keep 32-bit columns and gather from
    global memory while the rest use 12-bit halo indices.  FP64 steps match
    the oracle as in the default scenes; the distance-coded mixed form
    (its global-column blocks apply FP32 values) stays within 1e-6 of FP64
    in pressure."""
    dim, cfg, pos, vel, typ = H.build("dam3d_small", cg_tolerance=1e-9, influence_ratio_lap=8.0,
                                      max_neighbors=2560)
    say("wide halos, N=%d" % len(typ))
    g, o, recs = _compare_steps(lib3, cfg, dim, pos, vel, typ, 2)
    assert recs[0]["n_exact"]      # later steps start from positions with the CG's rounding
    for r in recs:
        assert r["type_exact"]
        assert r["p_gpu"] <= max(100 * r["p_spread"], 1e-9)
        assert r["x_gpu"] <= max(100 * r["x_spread"], 1e-13)
    p64 = np.array(g.scene.arr["pressure"])
    g.close()
    lib3.set_option("precision", 3)
    try:
        say("wide halos, coded mixed")
        gm = H.GpuRun(lib3, cfg, pos, vel, typ, n0=o.scene.ps.n0, lam=o.scene.ps.lambda_)
        for _ in range(2):
            assert gm.step() == 0
        pm = np.array(gm.scene.arr["pressure"])
        gm.close()
    finally:
        lib3.set_option("precision", 0)
    say("coded vs FP64 pressure", H.rel_err(pm, p64))
    assert H.rel_err(pm, p64) < 1e-6


def test_neighbour_overflow_is_reported(lib3):
    """max_neighbors 40 against ~110 neighbours per particle: the search
    returns MPS_ERR_NEIGHBOR_OVERFLOW and names the smallest original
    index among the particles with more than 40 neighbours."""
    dim, cfg, pos, vel, typ = H.build("dam3d_small")
    o = H.OracleRun(cfg, pos, vel, typ, dim)
    counts = np.array([len(x) for x in o.lists()])
    first = int(np.nonzero(counts > 40)[0][0])
    dim, cfg40, pos, vel, typ = H.build("dam3d_small", max_neighbors=40)
    lib3.set_option("exit_on_error", 0)
    try:
        lib3.set_config(cfg40)
        sc = A.HostScene(lib3.abi, pos, vel, typ)
        sc.ps.n0 = o.scene.ps.n0
        sc.ps.lambda_ = o.scene.ps.lambda_
        n = sc.num
        re_s = max(cfg40.influence_radius_lap, cfg40.influence_radius_n)
        dmin = (C.c_double * 3)(*[cfg40.domain_min[d] for d in range(3)])
        dmax = (C.c_double * 3)(*[cfg40.domain_max[d] for d in range(3)])
        nl = lib3.neighbor_list_create(n, 40)
        cl = lib3.cell_list_create(n, re_s, dmin, dmax)
        say("overflow search, N=%d" % n)
        lib3.neighbor_search_cell_linked_list(nl, sc.ptr(), cl, re_s)
        rc = lib3.mps_hip_last_error()
        msg = lib3.mps_hip_error_string(rc).decode()
        say("rc", rc, msg)
        assert rc == MPS_ERR_NEIGHBOR_OVERFLOW
        m = re.search(r"at particle (\d+)", msg)
        assert m and int(m.group(1)) == first
        lib3.mps_hip_release(sc.ptr())
        lib3.neighbor_list_free(nl)
        lib3.cell_list_free(cl)
    finally:
        lib3.set_option("exit_on_error", 1)


def test_nan_particle_becomes_ghost(lib3):
    """A fluid particle at NaN leaves the domain test and is turned into a
    ghost (boundary.c:33-61) by both the GPU and the oracle; it drops out
    of every later neighbour list, and the other particles keep the
    default tolerance."""
    dim, cfg, pos, vel, typ = H.build("dam3d_small", cg_tolerance=1e-10)
    i = int(np.nonzero(typ == 0)[0][len(np.nonzero(typ == 0)[0]) // 2])
    pos = pos.copy()
    pos[i] = np.nan
    say("NaN at particle", i)
    g, o, recs = _compare_steps(lib3, cfg, dim, pos, vel, typ, 2)
    assert g.scene.arr["type"][i] == GHOST and o.scene.arr["type"][i] == GHOST
    assert all(r["type_exact"] for r in recs)
    ok = np.ones(len(typ), bool)
    ok[i] = False
    a, b = g.scene.arr, o.scene.arr
    assert H.rel_err(a["pressure"][ok], b["pressure"][ok]) <= max(100 * recs[-1]["p_spread"], 1e-9)
    assert a["pressure"][i] == 0.0 and b["pressure"][i] == 0.0
    g.close()


def _isolated():
    """36 fluid particles 0.04-0.1 m apart inside the small dam break's
    tank (re = 0.031 m): each has no neighbour, so n = 0 (a surface
    particle) and the PPE has no unknowns."""
    dim, cfg, pos, vel, typ = H.build("dam3d_small")
    P = np.array([(x, y, z) for x in (0.05, 0.15, 0.25, 0.35) for y in (0.05, 0.15, 0.25)
                  for z in (0.02, 0.06, 0.10)], float)
    return dim, cfg, P, np.zeros_like(P), np.zeros(len(P), np.int32)


def test_empty_ppe(lib3):
    """No unknowns: the solve is skipped, pressure stays exactly 0, and
    the particles fall under gravity as in the oracle."""
    dim, cfg, pos, vel, typ = _isolated()
    dim0, cfg0, p0, v0, t0 = H.build("dam3d_small")
    n0 = H.OracleRun(cfg0, p0, v0, t0, dim0).scene.ps
    say("isolated particles", len(typ))
    lib3.set_option("tol_mode", 1)
    o = H.OracleRun(cfg, pos, vel, typ, dim, tol_mode=1, n0=n0.n0, lam=n0.lambda_)
    g = H.GpuRun(lib3, cfg, pos, vel, typ, n0=n0.n0, lam=n0.lambda_)
    for _ in range(3):
        assert g.step() == 0
        o.step()
    a, b = g.scene.arr, o.scene.arr
    assert np.all(a["pressure"] == 0.0) and np.all(b["pressure"] == 0.0)
    assert np.array_equal(a["on_surface"], b["on_surface"]) and np.all(a["on_surface"] == 1)
    assert np.array_equal(a["pos"].view(np.uint64), b["pos"].view(np.uint64))
    assert lib3.stats(g.scene).n_eq == 0
    g.close()


def _fields(sc):
    """The named fields of every particle, as bytes (not the struct's padding)."""
    return {f: np.ascontiguousarray(sc.arr[f]).tobytes() for f in H.FIELDS}


@pytest.mark.parametrize("precision", [0, 3])
def test_batch_with_empty_scene(lib3, precision):
    """A batch holding a scene with no unknowns next to an ordinary scene:
    both end with the bits of their single-scene runs, and the empty one
    reports no CG iterations."""
    from mps_amd import params
    lib3.set_option("tol_mode", 1)
    lib3.set_option("precision", precision)
    try:
        dim, cfg_e, pe, ve, te = _isolated()
        dim, cfg_d, pd, vd, td = H.build("dam3d_small", cg_tolerance=1e-10)
        o = H.OracleRun(cfg_d, pd, vd, td, dim)
        n0, lam = o.scene.ps.n0, o.scene.ps.lambda_

        def make(cfg, pos, vel, typ):
            lib3.set_config(cfg)
            sc = A.HostScene(lib3.abi, pos, vel, typ)
            sc.ps.n0 = n0
            sc.ps.lambda_ = lam
            return sc

        single = []
        for cfg, pos, vel, typ in ((cfg_e, pe, ve, te), (cfg_d, pd, vd, td)):
            sc = make(cfg, pos, vel, typ)
            lib3.check(lib3.mps_hip_run_steps(sc.ptr(), 3), "run_steps")
            lib3.check(lib3.mps_hip_download(sc.ptr(), None), "download")
            single.append((_fields(sc), lib3.stats(sc).total_iters))
            lib3.mps_hip_release(sc.ptr())
        scs = [make(cfg_e, pe, ve, te), make(cfg_d, pd, vd, td)]
        cfgs = [cfg_e, cfg_d]
        PSP = C.POINTER(lib3.abi.ParticleSystem)
        ps_arr = (PSP * 2)(*[C.pointer(s.ps) for s in scs])
        cf_arr = (C.POINTER(lib3.abi.SimConfig) * 2)(*[C.pointer(c) for c in cfgs])
        say("batch with an empty scene, precision", precision)
        lib3.check(lib3.mps_hip_run_batch(ps_arr, cf_arr, 2, 3), "run_batch")
        for q, sc in enumerate(scs):
            lib3.check(lib3.mps_hip_download(sc.ptr(), None), "download")
            assert _fields(sc) == single[q][0], q
            assert lib3.stats(sc).total_iters == single[q][1], q
            lib3.mps_hip_release(sc.ptr())
        assert single[0][1] == 0
    finally:
        lib3.set_option("precision", 0)


def test_finalize_then_batch_again(lib3):
    """mps_hip_finalize frees every context and the batched solvers'
    buffers; a batch run afterwards allocates them again and gives the
    bits of the same batch run before."""
    lib3.set_option("tol_mode", 1)
    lib3.set_option("precision", 3)
    try:
        dim, cfg_a, pa, va, ta = H.build("dam3d_small", cg_tolerance=1e-10)
        dim, cfg_b, pb, vb, tb = H.build("dam3d_small", cg_tolerance=1e-10, viscosity=1e-4)
        o = H.OracleRun(cfg_a, pa, va, ta, dim)
        n0, lam = o.scene.ps.n0, o.scene.ps.lambda_

        def run():
            scs = []
            for cfg, pos, vel, typ in ((cfg_a, pa, va, ta), (cfg_b, pb, vb, tb)):
                lib3.set_config(cfg)
                sc = A.HostScene(lib3.abi, pos, vel, typ)
                sc.ps.n0 = n0
                sc.ps.lambda_ = lam
                scs.append(sc)
            PSP = C.POINTER(lib3.abi.ParticleSystem)
            ps_arr = (PSP * 2)(*[C.pointer(s.ps) for s in scs])
            cf_arr = (C.POINTER(lib3.abi.SimConfig) * 2)(C.pointer(cfg_a), C.pointer(cfg_b))
            lib3.check(lib3.mps_hip_run_batch(ps_arr, cf_arr, 2, 2), "run_batch")
            out = []
            for sc in scs:
                lib3.check(lib3.mps_hip_download(sc.ptr(), None), "download")
                out.append(_fields(sc))
                lib3.mps_hip_release(sc.ptr())
            return out

        say("batch before finalize")
        first = run()
        lib3.mps_hip_finalize()
        say("batch after finalize")
        second = run()
        assert first == second
    finally:
        lib3.set_option("precision", 0)


MPS_ERR_SOLVER = 8


@pytest.mark.parametrize("dim,name", [(3, "shipped3d"), (2, "dam2d_l0.01")])
def test_iccg_wait_bound_reports_solver_error(request, dim, name):
    """Every wait of the IC(0) kernels is bounded by the option
    "ic_wait_limit".  With the bound at 1 poll the factor and the solves give
    up on whatever has not arrived at once, every kernel still runs to its
    end, and the step returns MPS_ERR_SOLVER; with the bound restored the
    next run is the normal one (factor bit-exact with the oracle)."""
    lib = request.getfixturevalue("lib3" if dim == 3 else "lib2")
    dim, cfg, pos, vel, typ = H.build(name, solver_type=1, cg_tolerance=1e-10)
    lib.set_option("tol_mode", 1)
    limit = lib.L.mps_hip_get_option(b"ic_wait_limit")
    assert limit >= 1024
    try:
        lib.set_option("ic_wait_limit", 1)
        g = H.GpuRun(lib, cfg, pos, vel, typ)
        rc = g.step()
        say("ICCG step with ic_wait_limit 1:", rc, lib.mps_hip_error_string(rc).decode())
        assert rc == MPS_ERR_SOLVER
        g.close()
    finally:
        lib.set_option("ic_wait_limit", limit)
    g = H.GpuRun(lib, cfg, pos, vel, typ)
    assert g.step() == 0
    g.close()


def test_brute_force_search_sets(lib3):
    """neighbor_search_brute_force (neighbor_search.c:44-81) on the GPU:
    every particle's list is the oracle's cell-search set, in ascending
    index order as the reference's O(N^2) loop lists it, on an off-lattice
    scene with ghosts (which list nothing and appear in no list)."""
    dim, cfg, pos, vel, typ = H.build("dam3d_small")
    rng = np.random.default_rng(5)
    pos = pos + rng.normal(0, 0.2 * 0.01, pos.shape)
    typ = typ.copy()
    typ[rng.choice(len(typ), 50, replace=False)] = 2
    o = H.OracleRun(cfg, pos, vel, typ, dim)
    want = o.lists()
    lib3.set_config(cfg)
    sc = A.HostScene(lib3.abi, pos, vel, typ)
    n = sc.num
    re_s = max(cfg.influence_radius_lap, cfg.influence_radius_n)
    nl = lib3.neighbor_list_create(n, cfg.max_neighbors)
    say("brute force search, N=%d" % n)
    lib3.neighbor_search_brute_force(nl, sc.ptr(), re_s)
    assert lib3.mps_hip_last_error() == 0
    got = A.padded_lists(nl.contents, n)
    for i in range(n):
        g = np.asarray(got[i])
        assert np.all(np.diff(g) > 0), i
        assert np.array_equal(g, np.sort(np.asarray(want[i]))), i
    assert all(len(got[i]) == 0 for i in np.nonzero(typ == 2)[0])
    lib3.mps_hip_release(sc.ptr())
    lib3.neighbor_list_free(nl)
