"""The CPU oracle against the reference compiled from its own sources
(oracle/_ref), bit for bit, over the scheme variants the reference offers.

Needs oracle/_ref (built where the reference sources exist); skipped
elsewhere -- tests/test_oracle_golden.py carries the same check through
stored fixtures.
"""
import ctypes as C
import ctypes.util

import numpy as np
import pytest

import scenes
from mps_amd import abi as A
from oracle.oracle import Oracle, Reference, make_config

pytestmark = pytest.mark.skipif(not Reference.available("3dp"), reason="reference build (oracle/_ref) absent")

FIELDS = ("pos", "vel", "acc", "pressure", "n", "type", "on_surface", "neighbor_count")


def _serial_openmp():
    # the reference's parallel n0 search breaks exact ties by thread order;
    # one thread gives its serial rule, which the oracle follows
    name = ctypes.util.find_library("gomp") or "libgomp.so.1"
    try:
        C.CDLL(name).omp_set_num_threads(1)
    except OSError:
        pass


def run_pair(scene, variant, steps, friction=False, final=None, **over):
    _serial_openmp()
    sc = scenes.CONFIGS[scene]
    dim = sc["dim"]
    p = scenes.params_for(scene, **over)
    cfg_o = make_config(A.types(dim, friction), **p)
    ref = Reference(variant)
    cfg_r = ref.config_from(cfg_o)
    ref.set_config(cfg_r)
    pos, vel, typ = scenes.dambreak(**sc)
    sr = A.HostScene(A.types(dim), pos, vel, typ)
    so = A.HostScene(A.types(dim), pos, vel, typ)
    ref.L.particle_system_calc_initial_params(sr.ptr())
    orc = Oracle(dim, friction)
    orc.init_params(so, cfg_o)
    for f in ("n0", "lambda_", "C_LL", "C_SL", "n0_count"):
        assert getattr(sr.ps, f) == getattr(so.ps, f), f
    nl, cl = ref.make_lists(sr.num)
    ref.L.neighbor_search_cell_linked_list(nl, sr.ptr(), cl, ref.search_radius())
    sim = orc.sim()
    sim.begin(so, cfg_o)
    for s in range(1, steps + 1):
        ref.L.simulation_step(sr.ptr(), nl, cl, s)
        sim.step(so, cfg_o)
        for f in FIELDS:
            assert np.ascontiguousarray(sr.arr[f]).tobytes() == np.ascontiguousarray(so.arr[f]).tobytes(), \
                (scene, variant, over, s, f)
    st, ix = sim.neighbor_lists()
    rl = A.padded_lists(nl.contents, sr.num)
    assert all(np.array_equal(rl[i], ix[st[i]:st[i + 1]]) for i in range(sr.num))
    ref.free_lists(nl, cl)
    if final is not None:
        final.update({f: np.array(so.arr[f]) for f in ("pos", "vel")})
    return sim.ppe(so.num)


VARIANTS_3D = [
    dict(),
    dict(clamp_negative_pressure=0),
    dict(collision_distance_ratio=0.0),
    dict(collision_distance_ratio=0.9, restitution_coeff=0.7),
    dict(cmps_gradient=1),
    dict(cmps_gradient=2),
    dict(surface_detection_method=1, surface_count_threshold=0.85),
    dict(ppe_type=1),
    dict(hs_mode=1, solver_type=1, influence_ratio_lap=2.1),
    dict(solver_type=1),
    dict(surface_tension_enabled=1, surface_tension_re_ratio=3.2, wetting_angle_SL=120.0),
    dict(use_analytical_lambda=1),
    dict(influence_ratio_n=3.1),
]


@pytest.mark.parametrize("over", VARIANTS_3D, ids=lambda d: ",".join("%s=%s" % kv for kv in d.items()) or "default")
def test_3d_parallel_variants(over):
    q = run_pair("shipped3d", "3dp", 4, cg_tolerance=1e-6, **over)
    assert q["n_eq"] > 0


def test_3d_serial_variant():
    run_pair("shipped3d", "3d", 4, cg_tolerance=1e-6, cmps_gradient=1)


FRICTION = dict(friction_enabled=1, friction_delta_theta=30.0, dynamic_friction_alpha=600.0,
                wetting_angle_SL=60.0)


@pytest.mark.parametrize("v_c", [0.01, 0.002])
def test_3d_serial_friction(v_c):
    """The hybrid wall friction (mps_fluid3D/src/operators.c:337-463) in the
    oracle's serial layout against the reference's own 3-D serial build,
    bit for bit over 12 steps of the shipped dam break, with the default
    switch speed v_c = 0.01 m/s and with 0.002 m/s, which move particles
    between the static and the dynamic branch.  The friction must change
    the result."""
    on, off = {}, {}
    run_pair("shipped3d", "3d", 12, friction=True, final=on, cg_tolerance=1e-6,
             friction_velocity_threshold=v_c, **FRICTION)
    run_pair("shipped3d", "3d", 12, friction=True, final=off, cg_tolerance=1e-6,
             friction_velocity_threshold=v_c, **dict(FRICTION, friction_enabled=0))
    changed = np.any(on["vel"] != off["vel"], axis=1)
    print("friction changed the velocity of %d particles" % changed.sum())
    assert changed.sum() > 0


@pytest.mark.parametrize("over", [dict(), dict(solver_type=1, hs_mode=1, influence_ratio_lap=4.0),
                                  dict(solver_type=1), dict(surface_tension_enabled=1, wetting_angle_SL=60.0),
                                  dict(collision_distance_ratio=0.0)],
                         ids=["cg", "iccg_hs", "iccg", "tension", "nocollide"])
def test_2d_variants(over):
    run_pair("dam2d_l0.01", "2d", 8, cg_tolerance=1e-6, **over)


def test_long_run_2d():
    """100 steps of the plain 2-D dam break stay bit-identical."""
    run_pair("dam2d_l0.01", "2d", 100, cg_tolerance=1e-6)


def test_brute_force_sets():
    """The reference's O(N^2) search gives the same sets (in index order) as
    the oracle's brute force and, as sets, as its cell search."""
    sc = scenes.CONFIGS["dam3d_small"]
    p = scenes.params_for("dam3d_small")
    cfg_o = make_config(A.types(3), **p)
    ref = Reference("3dp")
    ref.set_config(ref.config_from(cfg_o))
    pos, vel, typ = scenes.dambreak(**sc)
    rng = np.random.default_rng(5)
    pos = pos + rng.normal(0, 0.2 * sc["l0"], pos.shape)   # off-lattice
    typ = typ.copy()
    typ[rng.choice(len(typ), 50, replace=False)] = 2       # some ghosts
    s = A.HostScene(A.types(3), pos, vel, typ)
    n = s.num
    nl, cl = ref.make_lists(n)
    ref.L.neighbor_search_brute_force(nl, s.ptr(), ref.search_radius())
    brute = A.padded_lists(nl.contents, n)
    orc = Oracle(3)
    sim = orc.sim()
    sim.begin(s, cfg_o)
    st, ix = sim.neighbor_lists()
    for i in range(n):
        assert np.array_equal(np.sort(ix[st[i]:st[i + 1]]), brute[i])
    ref.free_lists(nl, cl)
