"""SimConfig for the benchmark workloads, built with the library's own
config_set_defaults (the reference's defaults, sim_config.c:20-79) and the
physics of the plain dam-break configurations: influence radii 3.1 l0
(Laplacian, gradient, PPE) and 2.1 l0 (number density), density-deviation
source (hs_mode 0, ppe_type 0), relaxation 0.2, surface threshold 0.97,
restitution 0.2 at 0.5 l0, negative-pressure clamp on, 3 wall + 2 dummy
layers."""
import ctypes as C


def bench_config(lib, w, solver=0, tol=1e-9, max_iter=20000):
    cfg = lib.abi.SimConfig()
    lib.config_set_defaults(C.byref(cfg))
    d = w["dims"]
    l0 = w["l0"]
    cfg.particle_distance = l0
    cfg.influence_ratio_lap = 3.1
    cfg.influence_ratio_n = 2.1
    cfg.max_neighbors = 512
    cfg.wall_layers = 3
    cfg.dummy_layers = 2
    cfg.density = 1000.0
    cfg.viscosity = w.get("nu", 1e-6)
    for k, g in enumerate((0.0, -9.81, 0.0)[:lib.dim]):
        cfg.gravity[k] = g
    cfg.dt = w["dt"]
    cfg.solver_type = solver
    cfg.cg_max_iter = max_iter
    cfg.cg_tolerance = tol
    cfg.relaxation_coeff = 0.2
    cfg.clamp_negative_pressure = 1
    cfg.ppe_type = 0
    cfg.hs_mode = 0
    cfg.surface_threshold = 0.97
    cfg.restitution_coeff = 0.2
    cfg.collision_distance_ratio = 0.5
    for k in range(lib.dim):
        cfg.domain_min[k] = 0.0
        cfg.domain_max[k] = d[k]
    cfg.influence_radius_lap = cfg.influence_ratio_lap * l0
    cfg.influence_radius_n = cfg.influence_ratio_n * l0
    cfg.influence_radius_st = cfg.surface_tension_re_ratio * l0
    return cfg


ENSEMBLE_SEED = 20260317


def ensemble(n=64, seed=ENSEMBLE_SEED, l0=0.01, dt=1e-3):
    """Independent 3-D dam breaks of uneven size: tank 0.8 (x) x 0.5 (y,
    vertical) x 0.4 (z) m; water column width ~ U[0.1, 0.4] m, height
    ~ U[0.2, 0.4] m, the full 0.4 m deep; kinematic viscosity log-uniform
    in [1e-6, 1e-4] m^2/s; l0 = 0.01 m, dt = 1e-3 s.  Seeded."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        w = float(rng.uniform(0.1, 0.4))
        h = float(rng.uniform(0.2, 0.4))
        nu = float(10.0 ** rng.uniform(-6.0, -4.0))
        out.append(dict(dims=(0.8, 0.5, 0.4, w, h, 0.4), l0=l0, dt=dt, nu=nu))
    return out
