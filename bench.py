#!/usr/bin/env python3
"""Throughput of the MPS timestep on one MI355X.

Workload (default): the plain 3-D dam break -- tank 1.0 (x) x 0.6 (y) x 0.5
(z) m, water column 0.3 x 0.5 x 0.5 m on a cubic lattice at l0 = 0.005 m with
3 wall + 2 dummy layers (1,077,505 particles, 594,000 fluid), rho = 1000,
nu = 1e-6, g = -9.81, dt = 5e-4, re = 3.1 l0 (Laplacian/gradient/PPE) and
2.1 l0 (number density), density-deviation PPE source, CG to relative
residual 1e-9, FP64.  Synthetic input: the lattice generator's particles, at
rest.

A step is one full timestep (viscosity/predictor, collision, counting sort
and neighbour search, density and flags, PPE assembly, CG solve, gradient
corrector, boundary) with all state resident in HBM.  W untimed warm-up
steps, then K timed steps between two device synchronisations.

Prints one JSON line.  `roofline` is the dominant kernel (the CG SpMV):
algorithmic bytes per launch / its mean duration measured with HIP events on
its stream during the timed steps.  `cpu_baseline` times the CPU oracle (the
C/OpenMP restatement of the reference, same stopping rule) on a bounded
sample of the same workload on this host.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import mps_amd  # noqa: E402
from mps_amd import abi as A  # noqa: E402

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "dam3d_l0.005": dict(dims=(1.0, 0.6, 0.5, 0.3, 0.5, 0.5), l0=0.005, dt=5e-4),
    "dam3d_l0.0035": dict(dims=(1.0, 0.6, 0.5, 0.3, 0.5, 0.5), l0=0.0035, dt=5e-4),
    "dam3d_l0.01": dict(dims=(1.0, 0.6, 0.5, 0.3, 0.5, 0.5), l0=0.01, dt=5e-4),
    # 64 seeded dam breaks of uneven size and viscosity (mps_amd.params.ensemble)
    "ensemble64": dict(ensemble=64),
    # the largest scene of that ensemble alone (column 0.4 x 0.4 x 0.4 m)
    "ensemble_scene": dict(dims=(0.8, 0.5, 0.4, 0.4, 0.4, 0.4), l0=0.01, dt=1e-3),
    # gradient of a 200-step 2-D dam break (tank 1.0 x 0.6, column 0.3 x 0.6)
    "adjoint2d": dict(adjoint=True, dims=(1.0, 0.6, 0.0, 0.3, 0.6, 0.0), l0=0.005, dt=5e-4, steps=200,
                      checkpoint=20, nu=1e-6, nu_target=5e-5, sigma_u0=0.01, seed=4),
}


def make_config(lib, w, solver, tol, max_iter=20000):
    from mps_amd.params import bench_config
    return bench_config(lib, w, solver, tol, max_iter)


def fluid_count(ps, n):
    """Fluid particles (type 0) of a ParticleSystem made by build_scene."""
    arr = np.ctypeslib.as_array(ps.contents.particles, shape=(n,))
    return int((arr["type"] == 0).sum())


def build_scene(lib, w):
    dims = (C.c_double * 6)(*w["dims"])
    n = lib.mps_gen_dambreak(dims, w["l0"], 3, 2, None, None)
    ps = lib.particle_system_create(n)
    got = lib.mps_gen_dambreak(dims, w["l0"], 3, 2, ps, None)
    if got != n:
        raise RuntimeError("scene generation failed")
    return ps, n


PRECISIONS = {"fp64": 0, "mixed": 1, "mixed32": 2, "mixed32q": 3}


def ic_solve_bytes(st, s=8):
    """Algorithmic bytes of one triangular solve (forward or backward): the
    strictly-lower entries (value + 4-byte column), 4 bytes of row offset
    per row, the right-hand side read and the result written."""
    return (st.nnz - st.n_eq) // 2 * (s + 4) + 4 * st.n_eq + 2 * s * st.n_eq


def coded_bytes(st):
    """Bytes of one application of the distance-coded FP32 operator
    (k_mx_spmv_q): one 4-byte word (halo index and distance code) per
    off-diagonal, and per row its length, FP64 Dirichlet diagonal, FP32 p
    read and FP32 Ap written.  Its own minimum, not the CSR figure."""
    z, ne = st.nnz, st.n_eq
    return 4 * (z - ne) + ne * (4 + 8 + 4 + 4)


def step_bytes(st, n, iters, s=8):
    """Algorithmic bytes of one whole timestep, B_particles + I * B_cg:
    six neighbour passes of 4 B per directed pair (search, density,
    viscosity, collision, assembly, p_min + gradient as two), the CSR and b
    written once, sort scatter and state traffic of about (22 s + 40) B per
    particle; per CG iteration the CSR once plus nine vector sweeps of s
    bytes per row."""
    e, z, ne = st.n_pairs, st.nnz, st.n_eq
    b_part = 28 * e + z * (s + 4) + 4 * (ne + 1) + 8 * ne + (22 * s + 40) * n
    b_cg = z * (s + 4) + 4 * (ne + 1) + 9 * s * ne
    return b_part + iters * b_cg


def spmv_bytes(st, s_val=8, s_vec=8):
    """Algorithmic bytes of one SpMV y = A x over the reference's CSR:
    Z (value + 4-byte column) + 4 (n_eq + 1) row pointers + x read once and
    y written once (Z = nonzeros with the diagonal).  The packed matrix the
    kernel reads stores 12-bit halo indices, so it moves fewer bytes than
    this."""
    z, ne = st.nnz, st.n_eq
    return z * (s_val + 4) + 4 * (ne + 1) + 2 * s_vec * ne


def cpu_baseline(w, tol, seconds_hint=30.0):
    """Oracle (C/OpenMP restatement of the reference) on this host: initial
    parameters, the initial search, then as many steps of the same workload
    as fit about 20 s (at least one)."""
    from oracle.oracle import Oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scenes  # the reference generator's lattice rules
    abi = A.types(3)
    orc = Oracle(3)
    cfg = make_config(mps_amd.load(3), w, 0, tol)
    d = w["dims"]
    pos, vel, typ = scenes.dambreak(3, d[:3], d[3:], w["l0"], 3, 2)
    sc = A.HostScene(abi, pos, vel, typ)
    orc.init_params(sc, cfg)
    sim = orc.sim(tol_mode=1)
    sim.begin(sc, cfg)
    steps, iters = 0, 0
    t0 = time.perf_counter()
    while True:
        sim.step(sc, cfg)
        steps += 1
        iters += sim.ppe(0)["iters"]
        el = time.perf_counter() - t0
        if el > 20.0 or steps >= 50:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=len(typ) * steps / el, unit="particle-steps/s", cores=threads, kind="port",
                sample="%d step(s) from rest of the %d-particle workload, CG relative 1e-9, "
                       "mean %.0f iterations/step, %.1f s" % (steps, len(typ), iters / steps, el))


class GpuBnorms:
    """sqrt(b.b) of the PPE step by step from rest, from a GPU run of the
    same scene; its right-hand side is bit-exact with the reference's."""

    def __init__(self, lib, w, tol, solver=0):
        self.lib = lib
        self.cfg = make_config(lib, w, solver, tol)
        lib.set_config(self.cfg)
        self.ps, _ = build_scene(lib, w)
        lib.particle_system_calc_initial_params(self.ps)
        self.iters = []

    def next(self):
        self.lib.set_config(self.cfg)
        self.lib.check(self.lib.mps_hip_run_steps(self.ps, 1), "step")
        st = self.lib.stats(self.ps)
        self.iters.append(st.last_iters)
        return st.last_bnorm

    def close(self):
        self.lib.mps_hip_release(self.ps)
        self.lib.particle_system_free(self.ps)


def cpu_baseline_reference(lib, w, tol, solver=0, seconds=20.0, max_steps=300):
    """The reference itself, compiled from its own sources with its OpenMP
    flags (oracle/_ref/libref3dp.so, built by oracle/Makefile), on this
    host: simulation_step of the same workload from rest, for as many steps
    as fit about `seconds` (at least one).  The reference's CG stops on an
    absolute residual (pressure_solver.c:90-131); step s gets tol * |b_s|,
    with |b_s| from a GPU run of the same steps kept in lockstep, so both
    stop on the same relative criterion.  Initial parameters come from the
    oracle (bit-identical to the reference's, whose O(N^2) pass takes hours
    at this size) and the initial search is not timed, as in
    simulation_run.  None when oracle/_ref is absent."""
    from oracle.oracle import Oracle, Reference
    if not Reference.available("3dp"):
        return None
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scenes  # the reference generator's lattice rules
    gb = GpuBnorms(lib, w, tol, solver)
    ref = Reference("3dp")
    cfg_o = make_config(lib, w, solver, tol)
    cfg_r = ref.config_from(cfg_o)
    d = w["dims"]
    pos, vel, typ = scenes.dambreak(3, d[:3], d[3:], w["l0"], 3, 2)
    so = A.HostScene(A.types(3), pos, vel, typ)
    Oracle(3).init_params(so, cfg_o)
    sr = A.HostScene(ref.abi, pos, vel, typ)
    for f in ("n0", "lambda_", "C_LL", "C_SL", "n0_count"):
        setattr(sr.ps, f, getattr(so.ps, f))
    ref.set_config(cfg_r)
    nl, cl = ref.make_lists(sr.num)
    ref.L.neighbor_search_cell_linked_list(nl, sr.ptr(), cl, ref.search_radius())
    steps, el = 0, 0.0
    for s in range(1, max_steps + 1):
        cfg_r.cg_tolerance = tol * gb.next()
        t0 = time.perf_counter()
        ref.L.simulation_step(sr.ptr(), nl, cl, s)
        el += time.perf_counter() - t0
        steps += 1
        if el > seconds:
            break
    ref.free_lists(nl, cl)
    gi = gb.iters
    gb.close()
    sname = "ICCG" if solver == 1 else "CG"
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=len(typ) * steps / el, unit="particle-steps/s", cores=threads, kind="reference",
                sample="%d step(s) from rest of the %d-particle workload through the reference's own "
                       "simulation_step (OpenMP, %d threads), %s to tol*|b| per step (the GPU's %s took "
                       "%.0f iterations/step on those steps), %.1f s"
                % (steps, len(typ), threads, sname, sname, sum(gi) / steps, el))


def ic_shape(lib):
    """(positions per chunk, waves per workgroup) of the IC(0) solve kernel
    the options select (16 waves only with 64-position chunks)."""
    ck = int(lib.L.mps_hip_get_option(b"ic_chunk"))
    w = int(lib.L.mps_hip_get_option(b"ic_waves"))
    return ck, (4 if w == 4 else (16 if w == 16 and ck == 64 else 8))


def profiled_traffic(precision, workload, nnz, name=None):
    """HBM bytes per launch of the solver's dominant kernel (the SpMV; the
    forward solve for ICCG, from profiles/iccg_<workload>.md), as measured
    by rocprofv3 (FETCH_SIZE + WRITE_SIZE scaled by the plain-copy
    calibration) and recorded in profiles/cg_<precision>_<workload>.md for a matrix of
    traffic_nnz nonzeros; scaled to this run's matrix by its nonzero count
    (the traffic is proportional to it).  None if no profile is recorded."""
    path = os.path.join(ROOT, "profiles", name or "cg_%s_%s.md" % (precision, workload))
    t = z = None
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                t = int(line.split("=")[1])
            elif line.startswith("traffic_nnz"):
                z = int(line.split("=")[1])
    except OSError:
        pass
    if t is None or not z:
        return None, None
    return int(round(t * nnz / z)), os.path.relpath(path, ROOT)


def run_ensemble(lib, a):
    """The ensemble workload: every scene resident, mps_hip_run_batch for
    W + K steps (K timed).  With --sequential the same scenes also run one
    after another through mps_hip_run_steps (the single-scene path), timed
    the same way, and the final states of the two runs are compared bit for
    bit."""
    from mps_amd import params
    ws = params.ensemble(WORKLOADS[a.workload]["ensemble"])

    def make():
        out = []
        for w in ws:
            cfg = params.bench_config(lib, w, 1 if a.solver == "iccg" else 0, a.tol)
            ps, n = build_scene(lib, w)
            lib.set_config(cfg)
            lib.particle_system_calc_initial_params(ps)
            out.append((cfg, ps, n))
        return out

    scenes = make()
    total = sum(m[2] for m in scenes)
    print("bench: %d scenes, %d particles (%d-%d per scene)" % (len(scenes), total, min(m[2] for m in scenes),
                                                                  max(m[2] for m in scenes)), file=sys.stderr, flush=True)
    tune = None
    if a.autotune and a.solver == "cg":
        # tiles of the batched CG from the largest scene
        big = max(scenes, key=lambda m: m[2])
        lib.set_config(big[0])
        res = mps_amd.TuneResult()
        lib.check(lib.mps_hip_autotune(big[1], 1, C.byref(res)), "autotune")
        tune = {"scene_particles": big[2], "spmv": res.spmv, "check_every": res.check_every,
                "spmv_ms": [round(v, 5) for v in res.spmv_ms], "solve_ms": [round(v, 2) for v in res.solve_ms],
                "pblk": res.pblk, "pblk_ms": [round(v, 3) for v in res.pblk_ms],
                "search": res.search, "search_ms": [round(v, 3) for v in res.search_ms]}
    PSP = C.POINTER(lib.abi.ParticleSystem)
    ps_arr = (PSP * len(scenes))(*[m[1] for m in scenes])
    cf_arr = (C.POINTER(lib.abi.SimConfig) * len(scenes))(*[C.pointer(m[0]) for m in scenes])
    if a.autotune and lib.mps_hip_get_option(b"batch_flat") == 1:
        # the particle kernels' block size for the launches over all scenes
        res = mps_amd.TuneResult()
        lib.check(lib.mps_hip_autotune_batch(ps_arr, cf_arr, len(scenes), 1, C.byref(res)), "autotune_batch")
        tune = dict(tune or {}, batch_pblk=res.pblk, batch_pblk_ms=[round(v, 3) for v in res.pblk_ms],
                    batch_search=res.search, batch_search_ms=[round(v, 3) for v in res.search_ms])
    lib.check(lib.mps_hip_run_batch(ps_arr, cf_arr, len(scenes), a.warmup), "warm-up steps")
    t0 = time.perf_counter()
    lib.check(lib.mps_hip_run_batch(ps_arr, cf_arr, len(scenes), a.steps), "timed steps")
    el = time.perf_counter() - t0
    iters = [lib.stats(m[1]).total_iters / max(lib.stats(m[1]).steps, 1) for m in scenes]
    seq = None
    if a.sequential:
        final = []
        for cfg, ps, n in scenes:
            lib.check(lib.mps_hip_download(ps, None), "download")
            final.append(_fields(lib, ps, n))
        seq_t = 0.0
        equal = True
        for q, (cfg, ps, n) in enumerate(make()):
            lib.set_config(cfg)
            lib.check(lib.mps_hip_run_steps(ps, a.warmup), "warm-up steps")
            t1 = time.perf_counter()
            lib.check(lib.mps_hip_run_steps(ps, a.steps), "timed steps")
            seq_t += time.perf_counter() - t1
            lib.check(lib.mps_hip_download(ps, None), "download")
            p1, x1 = _fields(lib, ps, n)
            equal = equal and p1.tobytes() == final[q][0].tobytes() and x1.tobytes() == final[q][1].tobytes()
            lib.mps_hip_release(ps)
            lib.particle_system_free(ps)
        seq = {"value": round(total * a.steps / seq_t, 1), "ms_per_step": round(1e3 * seq_t / a.steps, 3),
               "batch_speedup": round(seq_t / el, 3), "bitwise_equal_to_batch": bool(equal)}
    cpu = None
    if not a.no_cpu_baseline:
        try:
            cpu = cpu_baseline_reference(lib, ws[0], a.tol) if a.cpu_kind == "reference" else None
            if cpu is None:
                cpu = cpu_baseline(ws[0], a.tol)
            cpu["sample"] = "scene 0 of the ensemble: " + cpu["sample"]
        except Exception as e:  # baseline is reported, never required
            cpu = dict(error=str(e))
    print(json.dumps({
        "metric": "particle-steps/s, ensemble of independent 3D dam breaks",
        "value": round(total * a.steps / el, 1), "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * el / a.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic (reference dam-break lattice generator, seeded column sizes and viscosities)",
        "config": {"workload": a.workload, "scenes": len(scenes), "particles": total, "l0": ws[0]["l0"],
                   "dt": ws[0]["dt"], "seed": params.ENSEMBLE_SEED,
                   "tolerance": "relative %g" % a.tol, "mean_iters_per_step": round(sum(iters) / len(iters), 1),
                   "max_iters_per_step": round(max(iters), 1), "parallelism": "1 GPU, batched CG",
                   "precision": a.precision,
                   "graph": a.graph, "spmv_variant": int(lib.L.mps_hip_get_option(b"spmv")),
                   "check_every": int(lib.L.mps_hip_get_option(b"check_every"))},
        "autotune": tune, "sequential": seq, "cpu_baseline": cpu}), flush=True)
    for cfg, ps, n in scenes:
        lib.mps_hip_release(ps)
        lib.particle_system_free(ps)


def run_adjoint(a):
    """Gradient of L = mean |r_i(T) - r_i^target|^2 over the fluid (T = 200
    steps, checkpoints every 20) with respect to the viscosity, gravity and
    the initial velocities; the target is the same scene run with
    nu = 5e-5.  Times mps_hip_adjoint (forward with checkpoints, then the
    recomputed segments and the reverse sweep) and checks the gradient
    against central differences (relative step 1e-6; for nu 1e-6 of the
    target's viscosity): of the free forward, whose decisions flip at this
    horizon; of the frozen-decision replay, plain and Richardson-
    extrapolated; and, after one step, of the free forward."""
    w = WORKLOADS[a.workload]
    lib = mps_amd.load(2)
    lib.check(lib.mps_hip_init(0), "mps_hip_init")
    lib.set_option("tol_mode", 1)
    lib.set_option("exit_on_error", 0)
    from mps_amd.params import bench_config
    steps = w["steps"]

    def cfg_for(nu, g=None):
        c = bench_config(lib, dict(dims=w["dims"], l0=w["l0"], dt=w["dt"], nu=nu), 0, a.tol)
        if g is not None:
            c.gravity[0], c.gravity[1] = g
        return c

    dims = (C.c_double * 6)(*w["dims"])
    n = lib.mps_gen_dambreak(dims, w["l0"], 3, 2, None, None)
    base = lib.particle_system_create(n)
    lib.mps_gen_dambreak(dims, w["l0"], 3, 2, base, None)
    arr0 = np.ctypeslib.as_array(C.cast(base.contents.particles, C.POINTER(lib.abi.Particle)), shape=(n,))
    fl = arr0["type"] == 0
    u0 = np.zeros((n, 2))
    u0[fl] = np.random.default_rng(w["seed"]).normal(0.0, w["sigma_u0"], size=(fl.sum(), 2))
    cfg = cfg_for(w["nu"])
    lib.set_config(cfg)
    lib.particle_system_calc_initial_params(base)
    init = {f: getattr(base.contents, f) for f in ("n0", "lambda_", "C_LL", "C_SL", "n0_count")}

    def scene(u):
        q = lib.particle_system_create(n)
        lib.mps_gen_dambreak(dims, w["l0"], 3, 2, q, None)
        arr = np.ctypeslib.as_array(C.cast(q.contents.particles, C.POINTER(lib.abi.Particle)), shape=(n,))
        arr["vel"] = u
        for f, v in init.items():
            setattr(q.contents, f, v)
        return q, arr

    def final(nu=None, g=None, u=None):
        c = cfg_for(w["nu"] if nu is None else nu, g)
        lib.set_config(c)
        q, arr = scene(u0 if u is None else u)
        lib.check(lib.mps_hip_run_steps(q, steps), "forward")
        lib.check(lib.mps_hip_download(q, None), "download")
        out = (np.array(arr["pos"]), np.array(arr["type"]))
        lib.mps_hip_release(q)
        lib.particle_system_free(q)
        return out

    target = np.ascontiguousarray(final(nu=w["nu_target"])[0])

    def loss(**kw):
        pos, typ = final(**kw)
        f = typ == 0
        return float(np.sum((pos[f] - target[f]) ** 2) / f.sum())

    lib.set_config(cfg)
    q, _ = scene(u0)
    P = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    L, gnu, gg, gu = np.zeros(1), np.zeros(1), np.zeros(2), np.zeros((n, 2))
    t0 = time.perf_counter()
    lib.check(lib.mps_hip_adjoint(q, steps, w["checkpoint"], P(target), P(L), P(gnu), P(gg), P(gu)), "adjoint")
    el = time.perf_counter() - t0
    lib.particle_system_free(q)
    t1 = time.perf_counter()
    loss(); fwd = time.perf_counter() - t1
    rel = lambda x, y: abs(x - y) / max(abs(x), abs(y), 1e-300)
    if a.no_check:
        print(json.dumps({
            "metric": "adjoint gradient of a 200-step 2D dam break (forward + checkpointed reverse)",
            "value": round(el, 3), "unit": "s per gradient", "higher_is_better": False, "n_gpus": 1,
            "forward_only_s": round(fwd, 3), "gradient_over_forward": round(el / fwd, 2), "dtype": "f64",
            "config": {"workload": a.workload, "particles": n, "steps": steps, "checkpoint_every": w["checkpoint"]},
            "loss": L[0], "grad_nu": gnu[0], "grad_g": list(gg)}), flush=True)
        return
    checks = {}
    hnu = 1e-6 * w["nu_target"]
    checks["nu"] = rel(gnu[0], (loss(nu=w["nu"] + hnu) - loss(nu=w["nu"] - hnu)) / (2 * hnu))
    g = [0.0, -9.81]
    for d in range(2):
        gp, gm = list(g), list(g)
        gp[d] += 1e-6 * 9.81
        gm[d] -= 1e-6 * 9.81
        checks["g%d" % d] = rel(gg[d], (loss(g=gp) - loss(g=gm)) / (2e-6 * 9.81))
    v = np.zeros_like(u0)
    v[fl] = np.random.default_rng(w["seed"] + 1).normal(size=(fl.sum(), 2))
    hu = 1e-6 * np.linalg.norm(u0) / np.linalg.norm(v)
    checks["u0_direction"] = rel(float(np.sum(gu * v)), (loss(u=u0 + hu * v) - loss(u=u0 - hu * v)) / (2 * hu))
    # the same differences of the frozen-decision replay (the function the
    # adjoint differentiates), at relative steps h = 1e-6 and h/2: the plain
    # central difference at h, and Richardson's extrapolation
    # (4 D(h/2) - D(h)) / 3, which cancels the difference's own h^2 term
    hg = 1e-6 * 9.81
    sets = [(w["nu"], g, u0)]
    for f in (1.0, 0.5):
        sets += [(w["nu"] + f * hnu, g, u0), (w["nu"] - f * hnu, g, u0),
                 (w["nu"], [f * hg, -9.81], u0), (w["nu"], [-f * hg, -9.81], u0),
                 (w["nu"], [0.0, -9.81 + f * hg], u0), (w["nu"], [0.0, -9.81 - f * hg], u0),
                 (w["nu"], g, u0 + f * hu * v), (w["nu"], g, u0 - f * hu * v)]
    nus = np.array([x[0] for x in sets])
    gs = np.ascontiguousarray(np.array([x[1] for x in sets], dtype=np.float64))
    us = np.ascontiguousarray(np.array([x[2] for x in sets], dtype=np.float64))
    losses = np.zeros(len(sets))
    lib.set_config(cfg)
    q, _ = scene(u0)
    t2 = time.perf_counter()
    lib.check(lib.mps_hip_replay_losses(q, steps, P(target), len(sets), P(nus), P(gs), P(us), P(losses)), "replay")
    replay_s = time.perf_counter() - t2
    lib.particle_system_free(q)
    adj = {"nu": gnu[0], "g0": gg[0], "g1": gg[1], "u0_direction": float(np.sum(gu * v))}
    hs = {"nu": hnu, "g0": hg, "g1": hg, "u0_direction": hu}
    central, richardson = {}, {}
    for k, key in enumerate(adj):
        d1 = (losses[1 + 2 * k] - losses[2 + 2 * k]) / (2 * hs[key])
        d2 = (losses[9 + 2 * k] - losses[10 + 2 * k]) / hs[key]
        central[key] = rel(adj[key], d1)
        richardson[key] = rel(adj[key], (4 * d2 - d1) / 3)
    frozen = {"base_loss_vs_adjoint": rel(losses[0], L[0]), "central_h1e-6": central,
              "richardson_h1e-6": richardson}
    # one step: the free forward's decisions do not flip at these steps.  Its
    # own target (the nu_target run after one step, displaced by a seeded
    # 1 mm offset) keeps the loss at the size of the differences; the
    # viscosity step is 1e-4 of the target's, as a 1e-6 step moves the
    # one-step loss by ~1e-16 relative
    steps_saved, steps = steps, 1
    tpos, ttyp = final(nu=w["nu_target"])
    off = np.random.default_rng(w["seed"] + 2).normal(0.0, 1e-3, size=tpos.shape)
    target = np.ascontiguousarray(tpos + off * (ttyp == 0)[:, None])
    one = {}
    lib.set_config(cfg)
    q, _ = scene(u0)
    L1, gnu1, gg1, gu1 = np.zeros(1), np.zeros(1), np.zeros(2), np.zeros((n, 2))
    lib.check(lib.mps_hip_adjoint(q, 1, 1, P(target), P(L1), P(gnu1), P(gg1), P(gu1)), "adjoint 1 step")
    lib.particle_system_free(q)
    h1 = 1e-4 * w["nu_target"]
    one["nu"] = rel(gnu1[0], (loss(nu=w["nu"] + h1) - loss(nu=w["nu"] - h1)) / (2 * h1))
    for d in range(2):
        gp, gm = list(g), list(g)
        gp[d] += hg
        gm[d] -= hg
        one["g%d" % d] = rel(gg1[d], (loss(g=gp) - loss(g=gm)) / (2 * hg))
    one["u0_direction"] = rel(float(np.sum(gu1 * v)), (loss(u=u0 + hu * v) - loss(u=u0 - hu * v)) / (2 * hu))
    steps = steps_saved
    print(json.dumps({
        "metric": "adjoint gradient of a 200-step 2D dam break (forward + checkpointed reverse)",
        "value": round(el, 3), "unit": "s per gradient", "higher_is_better": False, "n_gpus": 1,
        "forward_only_s": round(fwd, 3), "gradient_over_forward": round(el / fwd, 2),
        "dtype": "f64", "data": "synthetic (reference dam-break lattice generator, seeded initial velocities)",
        "config": {"workload": a.workload, "particles": n, "fluid": int(fl.sum()), "steps": steps,
                   "checkpoint_every": w["checkpoint"], "nu": w["nu"], "nu_target": w["nu_target"],
                   "tolerance": "relative %g" % a.tol},
        "loss": L[0], "grad_nu": gnu[0], "grad_g": list(gg), "grad_u0_norm": float(np.linalg.norm(gu)),
        "fd_rel_err_one_step": one, "fd_rel_err_free_forward": checks, "fd_rel_err_frozen_decisions": frozen,
        "replay_s_for_%d_sets" % len(sets): round(replay_s, 3)}), flush=True)


def _fields(lib, ps, n):
    arr = np.ctypeslib.as_array(C.cast(ps.contents.particles, C.POINTER(lib.abi.Particle)), shape=(n,))
    return np.array(arr["pressure"]), np.array(arr["pos"])


def compare_forms(lib, ps, n, w, a, precision=None, emit=True):
    """The same scene, from the same initial parameters, resident for
    --steps steps in FP64 and in the mixed form: particle-steps/s of both
    and the relative L2 differences of pressure and positions at the end.
    Returns the result; prints it as one JSON line with emit."""
    precision = precision or a.precision
    mixed = PRECISIONS[precision] or 1
    init = {f: getattr(ps.contents, f) for f in ("n0", "lambda_", "C_LL", "C_SL", "n0_count")}
    out = {}
    for name, prec in (("fp64", 0), (precision if precision != "fp64" else "mixed", mixed)):
        q, _ = build_scene(lib, w)
        for f, v in init.items():
            setattr(q.contents, f, v)
        lib.set_option("precision", prec)
        lib.check(lib.mps_hip_run_steps(q, 1), "first step")
        t0 = time.perf_counter()
        lib.check(lib.mps_hip_run_steps(q, a.steps - 1), "steps")
        el = time.perf_counter() - t0
        lib.check(lib.mps_hip_download(q, None), "download")
        st = lib.stats(q)
        out[name] = dict(rate=n * (a.steps - 1) / el, ms=1e3 * el / (a.steps - 1),
                         iters=st.total_iters / max(st.steps, 1), fields=_fields(lib, q, n))
        print("bench: %s form, %d steps, %.3f ms/step" % (name, a.steps, out[name]["ms"]), file=sys.stderr,
              flush=True)
        lib.mps_hip_release(q)
        lib.particle_system_free(q)
    names = list(out)
    (p0, x0), (p1, x1) = out[names[0]]["fields"], out[names[1]]["fields"]
    rel = lambda u, v: float(np.linalg.norm(u - v) / max(np.linalg.norm(v), 1e-300))
    res = {"metric": "mixed vs FP64, same scene", "workload": a.workload, "particles": n,
           "steps": a.steps, "forms": {k: {"particle_steps_per_s": round(v["rate"], 1),
                                           "ms_per_step": round(v["ms"], 3),
                                           "mean_iters_per_step": round(v["iters"], 1)}
                                       for k, v in out.items()},
           "speedup": round(out[names[1]]["rate"] / out[names[0]]["rate"], 3),
           "rel_l2_pressure": rel(p1, p0), "rel_l2_position": rel(x1, x0)}
    lib.set_option("precision", 0)
    if emit:
        print(json.dumps(res), flush=True)
    return res


def sync_rate(lib, ps, n, cfg, a):
    """Throughput of the drop-in simulation_step (simulation.c:19-118) as
    the reference's own loop calls it: host ParticleSystem and padded
    NeighborList in, the step on the GPU, both back out."""
    dmin = (C.c_double * 3)(*[cfg.domain_min[d] for d in range(3)])
    dmax = (C.c_double * 3)(*[cfg.domain_max[d] for d in range(3)])
    re = max(cfg.influence_radius_lap, cfg.influence_radius_n)
    nl = lib.neighbor_list_create(n, cfg.max_neighbors)
    cl = lib.cell_list_create(n, re, dmin, dmax)
    lib.neighbor_search_cell_linked_list(nl, ps, cl, re)
    step = 0
    for _ in range(a.warmup):
        step += 1
        lib.simulation_step(ps, nl, cl, step)
    lib.check(lib.mps_hip_last_error(), "warm-up steps")
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step += 1
        lib.simulation_step(ps, nl, cl, step)
    el = time.perf_counter() - t0
    lib.check(lib.mps_hip_last_error(), "timed steps")
    st = lib.stats(ps)
    print(json.dumps({"metric": "particle-steps/s, drop-in simulation_step with host copies each step",
                      "value": round(n * a.steps / el, 1), "unit": "particle-steps/s",
                      "ms_per_step": round(1e3 * el / a.steps, 3), "steps": a.steps, "warmup": a.warmup,
                      "config": {"workload": a.workload, "particles": n, "max_neighbors": cfg.max_neighbors,
                                 "precision": a.precision, "last_iters": st.last_iters}}), flush=True)
    lib.neighbor_list_free(nl)
    lib.cell_list_free(cl)


def snapshot_rate(lib, w, a):
    """simulation_run (simulation.c:123-181) with CSV and VTK snapshots every
    a.snapshots steps, timed over the whole call, once with the snapshots
    written before the next step (async_output 0, the reference's order)
    and once written by a host thread while the device steps on
    (async_output 1).  Each run starts from a fresh scene; the files of the
    two runs are compared by SHA-1, then deleted."""
    import hashlib
    import shutil
    import tempfile
    res = {}
    digests = {}
    n = 0
    for mode in (0, 1):
        out = tempfile.mkdtemp(prefix="mps_snap_", dir=os.environ.get("TMPDIR", "/tmp"))
        cfg = make_config(lib, w, 0, a.tol)
        cfg.t_end = (a.steps + 0.5) * cfg.dt
        cfg.output_interval = a.snapshots
        cfg.output_dir = out.encode()
        lib.set_config(cfg)
        lib.set_option("async_output", mode)
        ps, n = build_scene(lib, w)
        lib.particle_system_calc_initial_params(ps)
        print("snapshots: async_output %d, %d steps, every %d" % (mode, a.steps, a.snapshots),
              file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        lib.simulation_run(ps)
        el = time.perf_counter() - t0
        lib.check(lib.mps_hip_last_error(), "simulation_run")
        lib.mps_hip_release(ps)
        lib.particle_system_free(ps)
        files = sorted(os.listdir(out))
        size = sum(os.path.getsize(os.path.join(out, f)) for f in files)
        digests[mode] = [(f, hashlib.sha1(open(os.path.join(out, f), "rb").read()).hexdigest())
                         for f in files]
        shutil.rmtree(out)
        res[mode] = {"seconds": round(el, 3), "particle_steps_per_s": round(n * a.steps / el, 1),
                     "files": len(files), "bytes": size}
    lib.set_option("async_output", 1)
    print(json.dumps({"metric": "particle-steps/s, simulation_run with CSV+VTK snapshots",
                      "value": res[1]["particle_steps_per_s"], "unit": "particle-steps/s",
                      "steps": a.steps, "snapshot_every": a.snapshots,
                      "sync_writer": res[0], "async_writer": res[1],
                      "speedup": round(res[0]["seconds"] / res[1]["seconds"], 3),
                      "files_identical": digests[0] == digests[1],
                      "config": {"workload": a.workload, "particles": n, "precision": a.precision}}),
          flush=True)


def calibrate_copy(nbytes=1 << 30, reps=5):
    """A plain copy of known size: each launch reads and writes exactly
    nbytes.  The HBM byte counters of its dispatches give the factor that
    turns the same counters on the solver's kernels into bytes."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps,
                      "GBps_read_plus_write": round(2 * nbytes * reps / el / 1e9, 1)}), flush=True)


def apply_options(lib, opts):
    """--option KEY=VALUE settings, applied after the defaults."""
    for kv in opts:
        key, _, val = kv.partition("=")
        lib.check(lib.mps_hip_set_option(key.encode(), float(val)), "option " + kv)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="dam3d_l0.005", choices=sorted(WORKLOADS))
    ap.add_argument("--solver", default="cg", choices=("cg", "iccg"))
    ap.add_argument("--precision", default="fp64", choices=tuple(PRECISIONS),
                    help="fp64; mixed = FP32 matrix, FP64 vectors; mixed32 = FP32 matrix and vectors")
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--spmv", type=int, default=0, help="SpMV kernel variant of the FP64 CG (0-2)")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: pick the SpMV variant and polling interval with mps_hip_autotune first (FP64 CG)")
    ap.add_argument("--graph", type=int, default=None,
                    help="1: replay the particle phases of each step from captured HIP graphs "
                         "(default: 1 for the ensemble workloads, 0 otherwise)")
    ap.add_argument("--fused", type=int, default=2,
                    help="FP64 CG iteration (same bits in every form): 2 three kernels (default), 0 five, 1 two")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mixed-check", default="mixed32q", choices=("none",) + tuple(p for p in PRECISIONS if p != "fp64"),
                    help="FP64 runs: also run this mixed form on the same scene and steps and report it beside")
    ap.add_argument("--cpu-kind", default="reference", choices=("reference", "port"),
                    help="CPU baseline: the reference built from its sources (oracle/_ref; the oracle "
                         "port where that is absent) or the oracle port")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-check", action="store_true",
                    help="adjoint workload: time the gradient only, without the finite-difference checks")
    ap.add_argument("--sync", action="store_true",
                    help="time the reference-style per-step drop-in calls (simulation_step with the "
                         "state and list copied in and out each step), PCIe included; not the headline")
    ap.add_argument("--compare", action="store_true",
                    help="run the scene in FP64 and in the chosen mixed form for --steps steps each and "
                         "report both rates and the relative L2 error of pressure and positions")
    ap.add_argument("--snapshots", type=int, default=0,
                    help="K > 0: time simulation_run over --steps steps with CSV+VTK snapshots every K "
                         "steps, with the synchronous and with the overlapped writer; not the headline")
    ap.add_argument("--sequential", action="store_true",
                    help="ensemble workloads: also time the scenes one after another and compare the bits")
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="set a library option (mps_hip_set_option) after the defaults; repeatable")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy 5 times (the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    a = ap.parse_args()
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (each box has one MI355X); "
              "--gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    w = WORKLOADS[a.workload]
    if w.get("adjoint"):
        run_adjoint(a)
        return
    lib = mps_amd.load(3)
    lib.check(lib.mps_hip_init(0), "mps_hip_init")
    if "ensemble" in w:
        if a.graph is None:
            a.graph = 1
        lib.set_option("tol_mode", 1)
        lib.set_option("exit_on_error", 0)
        lib.set_option("spmv", a.spmv)
        lib.set_option("graph", a.graph)
        lib.set_option("precision", PRECISIONS[a.precision])
        apply_options(lib, a.option)
        run_ensemble(lib, a)
        return
    if a.graph is None:
        a.graph = 0
    cfg = make_config(lib, w, 1 if a.solver == "iccg" else 0, a.tol)
    lib.set_config(cfg)
    lib.set_option("tol_mode", 1)
    lib.set_option("precision", PRECISIONS[a.precision])
    lib.set_option("exit_on_error", 0)
    lib.set_option("spmv", a.spmv)
    lib.set_option("fused", a.fused)
    lib.set_option("graph", a.graph)
    apply_options(lib, a.option)
    ps, n = build_scene(lib, w)
    print("bench: %s, %d particles, solver %s, %s" % (a.workload, n, a.solver, a.precision),
          file=sys.stderr, flush=True)
    lib.particle_system_calc_initial_params(ps)
    if a.compare:
        compare_forms(lib, ps, n, w, a)
        return
    if a.sync:
        sync_rate(lib, ps, n, cfg, a)
        return
    if a.snapshots > 0:
        lib.mps_hip_release(ps)
        lib.particle_system_free(ps)
        snapshot_rate(lib, w, a)
        return
    tune = None
    if a.autotune and a.precision == "fp64" and a.solver == "cg":
        res = mps_amd.TuneResult()
        t_tune = time.perf_counter()
        lib.check(lib.mps_hip_autotune(ps, 1, C.byref(res)), "autotune")
        tune = {"spmv": res.spmv, "check_every": res.check_every,
                "spmv_ms": [round(v, 5) for v in res.spmv_ms], "solve_ms": [round(v, 2) for v in res.solve_ms],
                "pblk": res.pblk, "pblk_ms": [round(v, 3) for v in res.pblk_ms],
                "search": res.search, "search_ms": [round(v, 3) for v in res.search_ms],
                "seconds": round(time.perf_counter() - t_tune, 2)}
    lib.check(lib.mps_hip_run_steps(ps, a.warmup), "warm-up steps")
    # the timed steps time only the dominant kernel's launches (two HIP
    # events each): the SpMV for CG, the forward triangular solve for ICCG
    iccg = a.solver == "iccg"
    lib.set_option("profile", 0 if a.no_profile else (3 if iccg else 2))
    lib.mps_hip_profile_reset()
    st0 = mps_amd.StepStats()
    lib.mps_hip_get_stats(ps, C.byref(st0))
    t0 = time.perf_counter()
    lib.check(lib.mps_hip_run_steps(ps, a.steps), "timed steps")   # returns after a device sync
    t1 = time.perf_counter()
    lib.set_option("profile", 0)
    st = mps_amd.StepStats()
    lib.mps_hip_get_stats(ps, C.byref(st))
    el = t1 - t0
    iters = (st.total_iters - st0.total_iters) / max(a.steps, 1)
    ms = C.c_double()
    cnt = C.c_longlong()
    lib.mps_hip_profile_get(b"ic_fwd" if iccg else b"spmv", C.byref(ms), C.byref(cnt))
    spmv_ms = ms.value / max(cnt.value, 1)
    spmv_launches = cnt.value
    roof = None
    if cnt.value > 0 and iccg:
        byt = ic_solve_bytes(st)
        ach = byt / (spmv_ms * 1e-3) / 1e9
        traffic, tsrc = profiled_traffic("fp64", a.workload, st.nnz, name="iccg_%s.md" % a.workload)
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=tsrc,
                    kernel="k_ic_csolve<0, false, %d, %d> (forward solve)" % ic_shape(lib), bytes_per_launch=int(byt),
                    mean_launch_ms=round(spmv_ms, 5), launches=spmv_launches)
    elif cnt.value > 0:
        byt = {"fp64": spmv_bytes(st, 8, 8), "mixed": spmv_bytes(st, 4, 8),
               "mixed32": spmv_bytes(st, 4, 4), "mixed32q": coded_bytes(st)}[a.precision]
        ach = byt / (spmv_ms * 1e-3) / 1e9
        traffic, tsrc = profiled_traffic(a.precision, a.workload, st.nnz)
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=tsrc,
                    kernel={"fp64": "k_spmv_pk", "mixed32q": "k_mx_spmv_q"}.get(a.precision, "k_mx_spmv_pk"), bytes_per_launch=int(byt),
                    mean_launch_ms=round(spmv_ms, 5), launches=spmv_launches)
    # per-kernel breakdown: one more step, every kernel class timed
    kernels = {}
    if not a.no_profile:
        lib.set_option("profile", 1)
        lib.mps_hip_profile_reset()
        lib.check(lib.mps_hip_run_steps(ps, 1), "profiled step")
        lib.set_option("profile", 0)
    for name in ("visc_predict", "collide_bin", "sort", "search", "assemble", "spmv", "axpy",
                 "pupdate", "grad_correct", "ic_factor", "ic_fwd", "ic_bwd"):
        if lib.mps_hip_profile_get(name.encode(), C.byref(ms), C.byref(cnt)) == 0 and cnt.value:
            kernels[name] = dict(total_ms=round(ms.value, 3), launches=cnt.value)
    lib.mps_hip_release(ps)
    step_roof = None
    if a.precision == "fp64" and a.solver == "cg":
        sb = step_bytes(st, n, iters)
        sach = sb / (el / a.steps) / 1e9
        step_roof = dict(bound="hbm", bytes_per_step=int(sb), iters_per_step=round(iters, 1),
                         achieved=round(sach, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(sach / HBM_PEAK_GBS, 4),
                         particle_steps_per_s_at_peak=round(n / (sb / (HBM_PEAK_GBS * 1e9)), 1))
    mixed = None
    if a.precision == "fp64" and a.solver == "cg" and a.mixed_check != "none":
        # the metric's other half: the mixed form on the same scene and steps
        r = compare_forms(lib, ps, n, w, a, precision=a.mixed_check, emit=False)
        fm = r["forms"]
        mixed = {"precision": a.mixed_check, "steps": a.steps,
                 "fp64_particle_steps_per_s": fm["fp64"]["particle_steps_per_s"],
                 "mixed_particle_steps_per_s": fm[a.mixed_check]["particle_steps_per_s"],
                 "speedup": r["speedup"], "rel_l2_pressure": r["rel_l2_pressure"],
                 "rel_l2_position": r["rel_l2_position"],
                 "mean_iters_per_step": {k: v["mean_iters_per_step"] for k, v in fm.items()}}
    cpu = None
    if not a.no_cpu_baseline and rank == 0:
        try:
            cpu = (cpu_baseline_reference(lib, w, a.tol, solver=1 if a.solver == "iccg" else 0)
                   if a.cpu_kind == "reference" else None)
            if cpu is None:
                cpu = cpu_baseline(w, a.tol)
        except Exception as e:  # baseline is reported, never required
            cpu = dict(error=str(e))
    value = n * a.steps / el
    line = {
        "metric": "particle-steps/s, 3D dam break 0.8M: FP64 vs mixed at <=1e-6 rel p error",
        "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * el / a.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic (reference dam-break lattice generator, fluid at rest)",
        "config": {"workload": a.workload, "particles": n, "fluid": fluid_count(ps, n), "l0": w["l0"],
                   "dt": w["dt"], "solver": a.solver, "precision": a.precision,
                   "tolerance": "relative %g" % a.tol, "mean_iters_per_step": round(iters, 1),
                   "n_eq": st.n_eq, "nnz": st.nnz, "pairs": st.n_pairs, "parallelism": "1 GPU",
                   "ic_levels": st.ic_levels if a.solver == "iccg" else None,
                   "spmv_variant": int(lib.L.mps_hip_get_option(b"spmv")),
                   "check_every": int(lib.L.mps_hip_get_option(b"check_every")), "graph": a.graph},
        "autotune": tune,
        "mixed": mixed,
        "roofline": roof,
        "step_roofline": step_roof,
        "cpu_baseline": cpu,
        "kernels_ms_one_step": kernels,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
