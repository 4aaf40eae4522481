"""Runs the same scene through the GPU library and the CPU oracle.

Both see identical inputs (positions from the reference generator's lattice,
the same SimConfig field values, the same n0/lambda) and the same solver
stopping rule, so their outputs can be compared field by field.
"""
import ctypes as C
import json
import os
import sys

import numpy as np

import scenes
from mps_amd import abi as A
from oracle.oracle import Oracle, make_config

FIELDS = ("pos", "vel", "acc", "pressure", "n", "type", "on_surface")
INIT_FIELDS = ("n0", "lambda_", "n0_count", "C_LL", "C_SL")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def friction_fixture(abi):
    """The wall-friction fixture (golden/friction3ds.npz, written by
    golden/make_friction_fixture.py from the reference's apply_friction):
    the scene in `abi`'s layout as it was before the friction, its config
    (3-D serial layout), the expected fluid velocities after, and the fluid
    mask."""
    g = np.load(os.path.join(GOLDEN, "friction3ds.npz"))
    params = json.loads(str(g["params"]))
    dim, cfg, pos, vel, typ = build(str(g["scene"]), friction=True, **params)
    typ = g["type"].astype(np.int32)
    fl = typ == 0
    pos[fl] = g["fluid_pos"]
    vel[fl] = g["fluid_vel_in"]
    sc = A.HostScene(abi, pos, vel, typ)
    sc.ps.n0 = float(g["n0"])
    sc.ps.lambda_ = float(g["lam"])
    return sc, cfg, np.array(g["fluid_vel_out"]), fl


def build(name, friction=False, **over):
    sc = scenes.CONFIGS[name]
    dim = sc["dim"]
    p = scenes.params_for(name, **over)
    cfg = make_config(A.types(dim, friction), **p)
    pos, vel, typ = scenes.dambreak(**sc)
    return dim, cfg, pos, vel, typ


class GpuRun:
    """Drives libmps_hip through its reference entry points (sync mode)."""

    def __init__(self, lib, cfg, pos, vel, typ, n0=None, lam=None):
        self.lib = lib
        self.cfg = cfg
        lib.set_config(cfg)
        self.scene = A.HostScene(lib.abi, pos, vel, typ)
        if n0 is None:
            lib.particle_system_calc_initial_params(self.scene.ptr())
        else:
            self.scene.ps.n0 = n0
            self.scene.ps.lambda_ = lam
        n = self.scene.num
        dim = lib.dim
        # the search radius of simulation.c:22-24 (re_st too with surface tension)
        re = max(cfg.influence_radius_lap, cfg.influence_radius_n)
        if cfg.surface_tension_enabled:
            re = max(re, cfg.influence_radius_st)
        dmin = (C.c_double * dim)(*[cfg.domain_min[d] for d in range(dim)])
        dmax = (C.c_double * dim)(*[cfg.domain_max[d] for d in range(dim)])
        self.nl = lib.neighbor_list_create(n, cfg.max_neighbors)
        self.cl = lib.cell_list_create(n, re, dmin, dmax)
        lib.neighbor_search_cell_linked_list(self.nl, self.scene.ptr(), self.cl, re)
        self.step_no = 0

    def step(self):
        self.lib.set_config(self.cfg)
        self.step_no += 1
        self.lib.simulation_step(self.scene.ptr(), self.nl, self.cl, self.step_no)
        rc = self.lib.mps_hip_last_error()
        return rc

    def lists(self):
        return A.padded_lists(self.nl.contents, self.scene.num)

    def close(self):
        self.lib.mps_hip_release(self.scene.ptr())
        self.lib.neighbor_list_free(self.nl)
        self.lib.cell_list_free(self.cl)


class OracleRun:
    def __init__(self, cfg, pos, vel, typ, dim, tol_mode=0, dot_order=0, n0=None, lam=None, friction=False):
        self.orc = Oracle(dim, friction)
        self.cfg = cfg
        self.scene = A.HostScene(A.types(dim), pos, vel, typ)
        if n0 is None:
            self.orc.init_params(self.scene, cfg)
        else:
            self.scene.ps.n0 = n0
            self.scene.ps.lambda_ = lam
        self.sim = self.orc.sim(tol_mode, dot_order)
        self.sim.begin(self.scene, cfg)

    def step(self):
        return self.sim.step(self.scene, self.cfg)

    def lists(self):
        st, ix = self.sim.neighbor_lists()
        return [ix[st[i]:st[i + 1]] for i in range(len(st) - 1)]

    def ppe(self):
        return self.sim.ppe(self.scene.num)


def rel_err(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    den = max(np.linalg.norm(b), 1e-300)
    return float(np.linalg.norm(a - b) / den)


def lists_equal(la, lb):
    if len(la) != len(lb):
        return False
    return all(np.array_equal(x, y) for x, y in zip(la, lb))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def step_compare(lib, name, steps, tol=1e-12, **over):
    """Runs `steps` steps of scene `name` (with config overrides) through the
    GPU library and two oracle runs that differ only in their dot-product
    order (forward and backward serial sums), all with relative CG
    tolerance `tol`.  Returns (gpu run, oracle run, per-step records) with
    the GPU's relative error against the oracle and the oracle's own
    dot-order spread for pressure, positions and velocities, and whether
    n, flags and types are bit-identical."""
    dim, cfg, pos, vel, typ = build(name, cg_tolerance=tol, **over)
    lib.set_option("tol_mode", 1)
    say("case", name, "N=%d" % len(typ), "steps", steps, over)
    o = OracleRun(cfg, pos, vel, typ, dim, tol_mode=1)
    ob = OracleRun(cfg, pos, vel, typ, dim, tol_mode=1, dot_order=1)
    g = GpuRun(lib, cfg, pos, vel, typ)
    # the GPU's own initial parameters (particle.c:69-221), bit for bit
    for f in INIT_FIELDS:
        assert getattr(g.scene.ps, f) == getattr(o.scene.ps, f), f
    out = []
    for s in range(1, steps + 1):
        assert g.step() == 0
        o.step()
        ob.step()
        out.append(compare_states(s, g.scene.arr, o.scene.arr, ob.scene.arr))
        say(out[-1])
    return g, o, out


def compare_states(step, a, b, c):
    """Record of GPU state a against oracle state b, with c the oracle run
    of the other dot-product order."""
    return dict(step=step,
                p_gpu=rel_err(a["pressure"], b["pressure"]),
                p_spread=rel_err(c["pressure"], b["pressure"]),
                x_gpu=rel_err(a["pos"], b["pos"]),
                x_spread=rel_err(c["pos"], b["pos"]),
                v_gpu=rel_err(a["vel"], b["vel"]),
                v_spread=rel_err(c["vel"], b["vel"]),
                n_exact=bool(np.array_equal(a["n"].view(np.uint64), b["n"].view(np.uint64))),
                flags_exact=bool(np.array_equal(a["on_surface"], b["on_surface"])),
                type_exact=bool(np.array_equal(a["type"], b["type"])))


def assert_ppe_exact(lib, g, o):
    """The GPU's last PPE, exported in the reference's layout, equals the
    oracle's bit for bit: unknown numbering, pattern, values, RHS."""
    q = lib.export_ppe(g.scene)
    oq = o.ppe()
    assert q["n_eq"] == oq["n_eq"]
    assert np.array_equal(q["eq_idx"], oq["eq_idx"])
    assert np.array_equal(q["row_ptr"], oq["row_ptr"].astype(np.int32))
    assert np.array_equal(q["col"], oq["col"])
    assert np.array_equal(q["val"].view(np.uint64), oq["val"].view(np.uint64))
    assert np.array_equal(q["b"].view(np.uint64), oq["b"].view(np.uint64))
    return q
