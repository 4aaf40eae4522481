"""GPU parity: libmps_hip (through its C-ABI) against the CPU oracle on the
same inputs.

Bit-exact: neighbour lists (set and order), number density, free-surface
flags, unknown numbering, PPE pattern, values and right-hand side, and the
initial n0/lambda.  Tolerance: pressure, velocity and position after the
CG solve, whose dot products sum in a different order than the reference's
serial loop.  Each tolerance is stated against the spread two oracle runs
show when only their dot-product order differs (forward vs backward serial
sums), which is the disagreement the reference itself would show.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.gpu


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_fp64_sqrt_div_match_host(lib3):
    """The kernels' FP64 sqrt and division round exactly like the host's."""
    rng = np.random.default_rng(7)
    n = 200000
    a = 10.0 ** rng.uniform(-12, 6, n)
    b = (10.0 ** rng.uniform(-12, 6, n)) * rng.choice([-1.0, 1.0], n)
    s = np.empty(n)
    q = np.empty(n)
    say("fp64 probe n=%d" % n)
    assert lib3.mps_hip_fp64_probe(n, _p(a), _p(b), _p(s), _p(q)) == 0
    assert np.array_equal(s.view(np.uint64), np.sqrt(a).view(np.uint64))
    assert np.array_equal(q.view(np.uint64), (a / b).view(np.uint64))


@pytest.mark.parametrize("name", ["shipped3d", "dam3d_small"])
def test_initial_params_bitexact(lib3, name):
    dim, cfg, pos, vel, typ = H.build(name)
    say("init params", name, len(typ))
    g = H.GpuRun(lib3, cfg, pos, vel, typ)
    o = H.OracleRun(cfg, pos, vel, typ, dim)
    assert g.scene.ps.n0 == o.scene.ps.n0
    assert g.scene.ps.lambda_ == o.scene.ps.lambda_
    # initial search: identical lists, in order
    assert H.lists_equal(g.lists(), o.lists())
    g.close()


def _step_compare(lib, name, steps, tol=1e-12, **over):
    return H.step_compare(lib, name, steps, tol, **over)


def test_one_step_3d_bitexact_parts(lib3):
    g, o, recs = _step_compare(lib3, "shipped3d", 1)
    r = recs[0]
    assert r["n_exact"] and r["flags_exact"] and r["type_exact"]
    assert H.lists_equal(g.lists(), o.lists())
    # PPE in the reference's layout
    H.assert_ppe_exact(lib3, g, o)
    # FP64 solve: within 100x the reference's own dot-order spread (or 1e-10)
    assert r["p_gpu"] <= max(100 * r["p_spread"], 1e-10)
    assert r["x_gpu"] <= max(100 * r["x_spread"], 1e-14)
    assert r["v_gpu"] <= max(100 * r["v_spread"], 1e-10)
    g.close()


def test_steps_3d_small(lib3):
    g, o, recs = _step_compare(lib3, "dam3d_small", 10)
    for r in recs:
        assert r["type_exact"]
        assert r["p_gpu"] <= max(100 * r["p_spread"], 1e-9)
        assert r["x_gpu"] <= max(100 * r["x_spread"], 1e-13)
    g.close()


def test_steps_2d(lib2):
    g, o, recs = _step_compare(lib2, "dam2d_l0.01", 10)
    assert recs[0]["n_exact"] and recs[0]["flags_exact"]
    for r in recs:
        assert r["type_exact"]
        assert r["p_gpu"] <= max(100 * r["p_spread"], 1e-9)
        assert r["x_gpu"] <= max(100 * r["x_spread"], 1e-13)
    g.close()


def test_hs_source_step(lib3):
    """High-order source: bit-exact right-hand side at step 1 (its inputs,
    the predicted velocities, are bit-exact then; from step 2 on they carry
    the CG's rounding)."""
    g, o, recs = _step_compare(lib3, "shipped3d", 1, hs_mode=1, influence_ratio_lap=2.1)
    assert recs[0]["n_exact"] and recs[0]["flags_exact"]
    q = lib3.export_ppe(g.scene)
    oq = o.ppe()
    assert np.array_equal(q["col"], oq["col"])
    assert np.array_equal(q["b"].view(np.uint64), oq["b"].view(np.uint64))
    assert recs[0]["p_gpu"] <= max(100 * recs[0]["p_spread"], 1e-10)
    g.close()


def test_resident_equals_sync(lib3):
    """mps_hip_run_steps (state resident in HBM) gives the same bits as the
    reference-style per-step calls that copy state in and out."""
    dim, cfg, pos, vel, typ = H.build("dam3d_small", cg_tolerance=1e-10)
    lib3.set_option("tol_mode", 1)
    g = H.GpuRun(lib3, cfg, pos, vel, typ)
    for _ in range(4):
        assert g.step() == 0
    from mps_amd import abi as A
    lib3.set_config(cfg)
    s2 = A.HostScene(lib3.abi, pos, vel, typ)
    s2.ps.n0 = g.scene.ps.n0
    s2.ps.lambda_ = g.scene.ps.lambda_
    say("resident run 4 steps")
    lib3.check(lib3.mps_hip_run_steps(s2.ptr(), 4), "run_steps")
    lib3.check(lib3.mps_hip_download(s2.ptr(), None), "download")
    for f in H.FIELDS:
        assert np.ascontiguousarray(g.scene.arr[f]).tobytes() == \
            np.ascontiguousarray(s2.arr[f]).tobytes(), f
    lib3.mps_hip_release(s2.ptr())
    g.close()


def _export_ic(lib, scene):
    st = lib.stats(scene)
    ne = st.n_eq
    lnnz = (st.nnz - ne) // 2
    rp = np.zeros(ne + 1, np.int32)
    col = np.zeros(max(lnnz, 1), np.int32)
    val = np.zeros(max(lnnz, 1))
    dinv = np.zeros(max(ne, 1))
    lib.check(lib.mps_hip_export_ic(scene.ptr(), _p(rp, C.c_int), _p(col, C.c_int), _p(val), _p(dinv)),
              "export_ic")
    return rp, col[:lnnz], val[:lnnz], dinv[:ne]


@pytest.mark.parametrize("dim,name", [(3, "shipped3d"), (2, "dam2d_l0.01")])
def test_iccg_exact_factor(lib3, lib2, dim, name):
    """ICCG: the IC(0) factor is the reference's, bit for bit (3-D sparse
    formulas; 2-D dense formulas), and the solve agrees to the dot-order
    spread."""
    lib = lib3 if dim == 3 else lib2
    g, o, recs = _step_compare(lib, name, 1, solver_type=1)
    r = recs[0]
    assert r["n_exact"] and r["flags_exact"]
    rp, col, val, dinv = _export_ic(lib, g.scene)
    oq = o.ppe()
    assert np.array_equal(rp, oq["l_ptr"].astype(np.int32))
    assert np.array_equal(col, oq["l_col"])
    assert np.array_equal(val.view(np.uint64), oq["l_val"].view(np.uint64))
    assert np.array_equal(dinv.view(np.uint64), oq["l_diag"].view(np.uint64))
    assert r["p_gpu"] <= max(100 * r["p_spread"], 1e-10)
    g.close()


@pytest.mark.parametrize("ratio", [3.2, 4.0])
def test_iccg_exact_factor_long_rows(lib3, ratio):
    """ICCG with PPE rows of more than 64 strictly-lower entries (Laplacian
    radius 3.2 and 4.0 l0: 73 and 125 in the lattice's interior, fewer at
    the walls and the surface), as a compressed flow also makes at 3.1 l0:
    the two-phase factor takes the long rows in the one-phase form, the
    short ones in its own, and the factor is the reference's bit for bit."""
    g, o, recs = _step_compare(lib3, "dam3d_small", 1, solver_type=1, influence_ratio_lap=ratio)
    r = recs[0]
    assert r["n_exact"] and r["flags_exact"]
    rp, col, val, dinv = _export_ic(lib3, g.scene)
    lower = np.diff(rp)
    say("ratio", ratio, "rows", len(lower), "longest strictly-lower row", int(lower.max()),
        "rows above 64:", int((lower > 64).sum()))
    assert lower.max() > 64 and (lower <= 64).any()
    oq = o.ppe()
    assert np.array_equal(rp, oq["l_ptr"].astype(np.int32))
    assert np.array_equal(col, oq["l_col"])
    assert np.array_equal(val.view(np.uint64), oq["l_val"].view(np.uint64))
    assert np.array_equal(dinv.view(np.uint64), oq["l_diag"].view(np.uint64))
    assert r["p_gpu"] <= max(100 * r["p_spread"], 1e-10)
    g.close()


@pytest.mark.parametrize("bucket,chunk", [(0, 32), (4, 64), (5, 128)])
def test_iccg_solve_orders(lib3, bucket, chunk):
    """The IC(0) triangular solves with other solve orders and chunk sizes
    (option "ic_bucket": levels per bucket 2^bucket, 0 giving the small
    dam break a bucket per level, several hundred of them; "ic_chunk" 64
    and 128, whose chunks can exceed the entries a chunk stages in LDS and
    then read their rows from HBM): pressure and positions within the
    oracle's dot-order spread over two steps."""
    saved = {k: lib3.L.mps_hip_get_option(k.encode()) for k in ("ic_bucket", "ic_chunk")}
    try:
        lib3.set_option("ic_bucket", bucket)
        lib3.set_option("ic_chunk", chunk)
        g, o, recs = _step_compare(lib3, "dam3d_small", 2, solver_type=1)
    finally:
        for k, v in saved.items():
            lib3.set_option(k, v)
    for r in recs:
        assert r["type_exact"]
        assert r["p_gpu"] <= max(100 * r["p_spread"], 1e-10)
        assert r["x_gpu"] <= max(100 * r["x_spread"], 1e-14)
    g.close()


def _laplacian(nx, ny, shift=0.05):
    n = nx * ny
    rows, cols, vals = [], [], []
    for i in range(nx):
        for j in range(ny):
            k = i * ny + j
            for di, dj in ((-1, 0), (0, -1), (0, 0), (0, 1), (1, 0)):
                a, b = i + di, j + dj
                if 0 <= a < nx and 0 <= b < ny:
                    rows.append(k)
                    cols.append(a * ny + b)
                    vals.append(4.0 + shift if (di, dj) == (0, 0) else -1.0)
    A = np.zeros((n, n))
    A[rows, cols] = vals
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, np.array(rows) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    order = np.lexsort((cols, rows))
    return A, rp, np.array(cols, np.int32)[order], np.array(vals)[order]


@pytest.mark.parametrize("solver", [0, 1])
def test_solve_csr_vs_numpy(lib3, solver):
    A, rp, col, val = _laplacian(24, 17)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(A.shape[0])
    x = np.zeros_like(b)
    it = C.c_int()
    lib3.set_option("tol_mode", 1)
    rc = lib3.mps_hip_solve_csr(A.shape[0], _p(rp, C.c_int), _p(col, C.c_int), _p(val), _p(b), _p(x),
                                solver, 2000, 1e-13, C.byref(it))
    assert rc == 0
    xs = np.linalg.solve(A, b)
    say("solver", solver, "iters", it.value, "err", H.rel_err(x, xs))
    assert H.rel_err(x, xs) < 1e-11
    assert 0 < it.value < 2000


def test_spmv_variants_same_bits(lib3):
    """Every SpMV variant (unroll depth, streaming loads, XCD block order)
    sums each row in the same order and the dot-product partials in the
    same block order, so the solves agree bit for bit."""
    dim, cfg, pos, vel, typ = H.build("dam3d_small", cg_tolerance=1e-10)
    lib3.set_option("tol_mode", 1)
    out = []
    try:
        for v in range(3):
            lib3.set_option("spmv", v)
            say("spmv variant", v)
            g = H.GpuRun(lib3, cfg, pos, vel, typ)
            for _ in range(2):
                assert g.step() == 0
            out.append(np.ascontiguousarray(g.scene.arr["pressure"]).tobytes())
            g.close()
    finally:
        lib3.set_option("spmv", 0)
    assert all(o == out[0] for o in out)


@pytest.mark.parametrize("precision", [0, 1, 2, 3])
def test_fused_cg_same_bits(lib3, precision):
    """The fused CG iterations give the five-kernel form's bits.  FP64: two
    kernels (previous iteration closed inside the SpMV, p formed where it
    is read) and three (each finaliser's reduction repeated by the blocks
    of the next vector kernel).  Mixed forms: three kernels, the
    reliable-update fold and the stop tests repeated the same way."""
    dim, cfg, pos, vel, typ = H.build("dam3d_mid", cg_tolerance=1e-10)
    lib3.set_option("tol_mode", 1)
    lib3.set_option("precision", precision)
    forms = (0, 1, 2) if precision == 0 else (0, 2)
    out = []
    try:
        for f in forms:
            lib3.set_option("fused", f)
            say("precision", precision, "fused", f)
            g = H.GpuRun(lib3, cfg, pos, vel, typ)
            for _ in range(2):
                assert g.step() == 0
            st = lib3.stats(g.scene)
            out.append((np.ascontiguousarray(g.scene.arr["pressure"]).tobytes(), st.last_iters,
                        st.last_residual, st.last_outer, st.total_iters))
            g.close()
    finally:
        lib3.set_option("fused", 2)
        lib3.set_option("precision", 0)
    assert all(o == out[0] for o in out)


@pytest.mark.parametrize("name,libname", [("dam3d_mid", "lib3"), ("dam2d_l0.01", "lib2")])
def test_graph_same_bits(request, name, libname):
    """The particle phases replayed from captured HIP graphs (option
    "graph") leave every field with the bits of plain launches, over steps
    of both buffer parities."""
    lib = request.getfixturevalue(libname)
    dim, cfg, pos, vel, typ = H.build(name, cg_tolerance=1e-10)
    lib.set_option("tol_mode", 1)
    out = []
    try:
        for gmode in (0, 1):
            lib.set_option("graph", gmode)
            say("graph", gmode, name)
            g = H.GpuRun(lib, cfg, pos, vel, typ)
            lib.mps_hip_set_resident(g.scene.ptr(), 1)
            assert lib.mps_hip_run_steps(g.scene.ptr(), 5) == 0
            assert lib.mps_hip_download(g.scene.ptr(), None) == 0
            st = lib.stats(g.scene)
            out.append(([np.ascontiguousarray(g.scene.arr[f]).tobytes() for f in H.FIELDS], st.total_iters))
            g.close()
    finally:
        lib.set_option("graph", 0)
    assert out[0][1] == out[1][1]
    for f, a, b in zip(H.FIELDS, out[0][0], out[1][0]):
        assert a == b, f


def test_autotune_keeps_bits(lib3):
    """mps_hip_autotune times the SpMV variants and polling intervals on a
    private copy of the scene, leaves the caller's particles as they were,
    applies the fastest choices, and the steps that follow give the bits of
    the default options."""
    import mps_amd
    dim, cfg, pos, vel, typ = H.build("dam3d_mid", cg_tolerance=1e-10)
    lib3.set_option("tol_mode", 1)
    out = []
    try:
        for tune in (False, True):
            g = H.GpuRun(lib3, cfg, pos, vel, typ)
            if tune:
                names = g.scene.arr.dtype.names
                before = [np.ascontiguousarray(g.scene.arr[f]).tobytes() for f in names]
                res = mps_amd.TuneResult()
                assert lib3.mps_hip_autotune(g.scene.ptr(), 1, C.byref(res)) == 0
                say("autotune: spmv %d (%s ms), check_every %d (%s ms), %d iterations"
                    % (res.spmv, ["%.4f" % v for v in res.spmv_ms], res.check_every,
                       ["%.2f" % v for v in res.solve_ms], res.iters))
                assert [np.ascontiguousarray(g.scene.arr[f]).tobytes() for f in names] == before
                assert res.spmv in (0, 1) and res.check_every in (8, 16, 32, 64)
                assert res.spmv_ms[res.spmv] == min(res.spmv_ms[0], res.spmv_ms[1])
                assert res.solve_ms[(8, 16, 32, 64).index(res.check_every)] == min(res.solve_ms)
                assert min(res.spmv_ms) > 0 and min(res.solve_ms) > 0 and res.iters > 0
                assert lib3.L.mps_hip_get_option(b"spmv") == res.spmv
                assert lib3.L.mps_hip_get_option(b"check_every") == res.check_every
                say("autotune: pblk %d (%s ms)" % (res.pblk, ["%.3f" % v for v in res.pblk_ms]))
                assert res.pblk in (128, 256, 512) and min(res.pblk_ms) > 0
                assert res.pblk_ms[(128, 256, 512).index(res.pblk)] == min(res.pblk_ms)
                assert lib3.L.mps_hip_get_option(b"pblk") == res.pblk
                assert res.search in (1, 2) and min(res.search_ms) > 0
                assert res.search_ms[res.search - 1] == min(res.search_ms)
                assert lib3.L.mps_hip_get_option(b"search") == res.search
            for _ in range(2):
                assert g.step() == 0
            out.append([np.ascontiguousarray(g.scene.arr[f]).tobytes() for f in H.FIELDS])
            g.close()
    finally:
        lib3.set_option("spmv", 0)
        lib3.set_option("check_every", 16)
        lib3.set_option("pblk", 256)
        lib3.set_option("search", 1)
    for f, a, b in zip(H.FIELDS, out[0], out[1]):
        assert a == b, f


@pytest.mark.parametrize("graph", [0, 1])
def test_particle_block_sizes_same_bits(lib3, graph):
    """The per-particle kernels at block sizes 128, 256 and 512 (option
    "pblk") give the same bits over three steps, with plain launches and
    with captured graphs (a new size drops the captures); sizes other
    than these are refused."""
    dim, cfg, pos, vel, typ = H.build("dam3d_small", cg_tolerance=1e-10)
    lib3.set_option("tol_mode", 1)
    lib3.set_option("graph", graph)
    out = []
    try:
        assert lib3.mps_hip_set_option(b"pblk", 192.0) != 0
        for pb in (256, 128, 512):
            lib3.set_option("pblk", pb)
            say("pblk", pb, "graph", graph)
            g = H.GpuRun(lib3, cfg, pos, vel, typ)
            lib3.check(lib3.mps_hip_set_resident(g.scene.ptr(), 1), "resident")
            lib3.check(lib3.mps_hip_run_steps(g.scene.ptr(), 3), "run_steps")
            lib3.check(lib3.mps_hip_download(g.scene.ptr(), None), "download")
            out.append([np.ascontiguousarray(g.scene.arr[f]).tobytes() for f in H.FIELDS])
            g.close()
    finally:
        lib3.set_option("pblk", 256)
        lib3.set_option("graph", 0)
    for o in out[1:]:
        for f, a, b in zip(H.FIELDS, out[0], o):
            assert a == b, f


@pytest.mark.parametrize("name,libname", [("dam3d_mid", "lib3"), ("dam2d_l0.01", "lib2")])
def test_search_forms_same_bits(request, name, libname):
    """The neighbour search with candidates staged in LDS (option "search"
    1, the default, and 2, reading two candidates at a time) and gathered
    from HBM (0) give the same lists, in order, and the same bits in every
    field over three steps, with plain launches and with captured graphs;
    other values are refused."""
    lib = request.getfixturevalue(libname)
    dim, cfg, pos, vel, typ = H.build(name, cg_tolerance=1e-10)
    lib.set_option("tol_mode", 1)
    out = []
    try:
        assert lib.mps_hip_set_option(b"search", 3.0) != 0
        for graph in (0, 1):
            lib.set_option("graph", graph)
            for form in (1, 2, 0):
                lib.set_option("search", form)
                say(name, "search form", form, "graph", graph)
                g = H.GpuRun(lib, cfg, pos, vel, typ)
                lists0 = g.lists()
                lib.check(lib.mps_hip_set_resident(g.scene.ptr(), 1), "resident")
                lib.check(lib.mps_hip_run_steps(g.scene.ptr(), 3), "run_steps")
                lib.check(lib.mps_hip_download(g.scene.ptr(), g.nl), "download")
                out.append((lists0, g.lists(),
                            [np.ascontiguousarray(g.scene.arr[f]).tobytes() for f in H.FIELDS]))
                g.close()
    finally:
        lib.set_option("search", 1)
        lib.set_option("graph", 0)
    for o in out[1:]:
        assert H.lists_equal(out[0][0], o[0]), "initial lists"
        assert H.lists_equal(out[0][1], o[1]), "lists after three steps"
        for f, a, b in zip(H.FIELDS, out[0][2], o[2]):
            assert a == b, f


@pytest.mark.parametrize("flat,graph", [(1, 0), (1, 1), (0, 0)])
@pytest.mark.parametrize("precision", [0, 1, 2, 3])
def test_batch_equals_single_runs(lib3, precision, flat, graph):
    """mps_hip_run_batch leaves every scene with the bits and CG iteration
    counts of its own single-scene run (scenes of different sizes and
    viscosities): with one launch per operator for all scenes (option
    "batch_flat" 1, the default), also with the step replayed from
    captured graphs (option "graph": the phases before and after the solve,
    and a chunk of CG iterations), and with per-scene launches (0); one
    batched CG for all scenes in every form: FP64, or the mixed forms of
    precisions 1-3."""
    lib3.set_option("precision", precision)
    lib3.set_option("batch_flat", flat)
    try:
        _batch_equals_single(lib3, graph=graph)
    finally:
        lib3.set_option("precision", 0)
        lib3.set_option("batch_flat", 1)
        lib3.set_option("graph", 0)


def test_autotune_batch_keeps_bits(lib3):
    """mps_hip_autotune_batch times the batch's particle phases at each
    block size on private copies of the scenes and picks one; the scenes
    it was given are left untouched, and the batch run with its choice has
    the bits of the single-scene runs."""
    try:
        _batch_equals_single(lib3, tune=True)
    finally:
        lib3.set_option("pblk", 256)
        lib3.set_option("search", 1)


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("precision", [0, 3])
def test_batch_equals_single_runs_2d(lib2, precision, graph):
    """The 2-D library's batches (particle phases per scene, one batched
    CG, its iterations replayed from a captured chunk with option "graph"):
    every scene keeps the bits and CG iteration counts of its own run."""
    lib2.set_option("precision", precision)
    try:
        _batch_equals_single(lib2, graph=graph, scenes_w=SCENES_2D)
    finally:
        lib2.set_option("precision", 0)
        lib2.set_option("graph", 0)


def test_batch_after_single_steps(lib3):
    """Scenes that took different numbers of single-scene steps first (so
    their buffer halves differ) then run batched: each ends with the bits
    of the same number of steps taken alone."""
    _batch_equals_single(lib3, pre=(1, 0, 2))


SCENES_3D = [dict(dims=(0.4, 0.3, 0.12, 0.12, 0.2, 0.12), l0=0.01, dt=5e-4, nu=1e-6),
             dict(dims=(0.6, 0.4, 0.2, 0.2, 0.3, 0.2), l0=0.01, dt=5e-4, nu=3e-5),
             dict(dims=(0.4, 0.3, 0.12, 0.2, 0.15, 0.12), l0=0.01, dt=5e-4, nu=1e-4)]
SCENES_2D = [dict(dims=(0.6, 0.4, 0.0, 0.15, 0.3, 0.0), l0=0.005, dt=5e-4, nu=1e-6),
             dict(dims=(1.0, 0.6, 0.0, 0.3, 0.5, 0.0), l0=0.01, dt=5e-4, nu=3e-5),
             dict(dims=(0.8, 0.5, 0.0, 0.25, 0.25, 0.0), l0=0.005, dt=5e-4, nu=1e-4)]


def _batch_equals_single(lib3, pre=(0, 0, 0), graph=0, tune=False, scenes_w=SCENES_3D):
    import ctypes as C
    from mps_amd import params
    steps = 4
    lib3.set_option("tol_mode", 1)

    def make(w):
        cfg = params.bench_config(lib3, w, 0, 1e-10, 20000)
        dims = (C.c_double * 6)(*w["dims"])
        n = lib3.mps_gen_dambreak(dims, w["l0"], 3, 2, None, None)
        ps = lib3.particle_system_create(n)
        assert lib3.mps_gen_dambreak(dims, w["l0"], 3, 2, ps, None) == n
        lib3.set_config(cfg)
        lib3.particle_system_calc_initial_params(ps)
        return cfg, ps, n

    def fields(ps, n):
        arr = np.ctypeslib.as_array(C.cast(ps.contents.particles, C.POINTER(lib3.abi.Particle)), shape=(n,))
        return {f: np.ascontiguousarray(arr[f]).tobytes() for f in H.FIELDS}

    single = []
    for w in scenes_w:
        cfg, ps, n = make(w)
        lib3.set_config(cfg)
        say("single scene", n, "nu", w["nu"])
        lib3.check(lib3.mps_hip_run_steps(ps, steps + pre[len(single)]), "run_steps")
        lib3.check(lib3.mps_hip_download(ps, None), "download")
        single.append((fields(ps, n), lib3.stats(ps).total_iters))
        lib3.mps_hip_release(ps)
        lib3.particle_system_free(ps)
    made = [make(w) for w in scenes_w]
    for (cfg, ps, n), k in zip(made, pre):
        if k:
            lib3.set_config(cfg)
            lib3.check(lib3.mps_hip_run_steps(ps, k), "steps before the batch")
    cfgs = [m[0] for m in made]
    PSP = C.POINTER(lib3.abi.ParticleSystem)
    ps_arr = (PSP * len(made))(*[m[1] for m in made])
    cf_arr = (C.POINTER(lib3.abi.SimConfig) * len(made))(*[C.pointer(c) for c in cfgs])
    if tune:
        import mps_amd
        res = mps_amd.TuneResult()
        lib3.check(lib3.mps_hip_autotune_batch(ps_arr, cf_arr, len(made), 1, C.byref(res)), "autotune_batch")
        say("batch tuner: pblk", res.pblk, "ms", list(res.pblk_ms), "search", res.search, list(res.search_ms))
        assert res.pblk in (128, 256, 512) and lib3.mps_hip_get_option(b"pblk") == res.pblk
        assert all(v > 0 for v in res.pblk_ms)
        assert res.search in (1, 2) and lib3.mps_hip_get_option(b"search") == res.search
        assert res.search_ms[res.search - 1] == min(res.search_ms)
    say("batch of", len(made), "graph", graph)
    lib3.set_option("graph", graph)
    try:
        # in two calls (each captures anew); the second call's third step
        # replays the capture of its first
        lib3.check(lib3.mps_hip_run_batch(ps_arr, cf_arr, len(made), 1), "run_batch")
        lib3.check(lib3.mps_hip_run_batch(ps_arr, cf_arr, len(made), steps - 1), "run_batch")
    finally:
        lib3.set_option("graph", 0)
    for q, (cfg, ps, n) in enumerate(made):
        lib3.check(lib3.mps_hip_download(ps, None), "download")
        got = fields(ps, n)
        for f in H.FIELDS:
            assert got[f] == single[q][0][f], (q, f)
        assert lib3.stats(ps).steps == steps + pre[q]
        assert lib3.stats(ps).total_iters == single[q][1]
        lib3.mps_hip_release(ps)
        lib3.particle_system_free(ps)
