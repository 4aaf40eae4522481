"""Synthetic scenes for tests and the benchmark.

`dambreak` follows the lattice and typing rules of the reference generator
(mps_fluid3D_parallel/tools/gen_dambreak.py:99-166; 2-D:
mps_fluid2D/tools/gen_dambreak.py:107-150): sites in i->j->k order (x
slowest), wall where an index is on or past a tank face (no lid), dummy
where it is wall_layers or more past it, fluid inside the column.  Positions
carry the %.8e rounding of the particles.txt text the reference reads.
"""
import numpy as np

# Scene geometries of the benchmark configurations (x, y vertical, z).
CONFIGS = {
    # plain 2-D dam break: tank 1.0 x 0.6, column 0.25 x 0.5, l0 0.01
    "dam2d_l0.01": dict(dim=2, domain=(1.0, 0.6), water=(0.25, 0.5), l0=0.01, wl=3, dl=2,
                        dt=1e-3),
    # plain 3-D dam break: tank 1.0 (x) x 0.6 (y) x 0.5 (z), column 0.3 x 0.5 x 0.5, l0 0.005
    "dam3d_l0.005": dict(dim=3, domain=(1.0, 0.6, 0.5), water=(0.3, 0.5, 0.5), l0=0.005,
                         wl=3, dl=2, dt=5e-4),
    # the same refined to l0 0.0035
    "dam3d_l0.0035": dict(dim=3, domain=(1.0, 0.6, 0.5), water=(0.3, 0.5, 0.5), l0=0.0035,
                          wl=3, dl=2, dt=5e-4),
    # 2-D column 0.3 x 0.6 at l0 0.005 (differentiable case)
    "dam2d_l0.005": dict(dim=2, domain=(1.0, 0.6), water=(0.3, 0.6), l0=0.005, wl=3, dl=2,
                         dt=5e-4),
    # geometry of the reference's shipped 3-D example (28,112 particles)
    "shipped3d": dict(dim=3, domain=(0.584, 0.292, 0.024), water=(0.146, 0.292, 0.024),
                      l0=0.008, wl=2, dl=2, dt=5e-4),
    # geometry of the reference's shipped 2-D example (1,928 particles)
    "shipped2d": dict(dim=2, domain=(0.584, 0.292), water=(0.146, 0.292), l0=0.008, wl=4,
                      dl=4, dt=5e-4),
    # small 3-D scenes for quick parity runs
    "dam3d_small": dict(dim=3, domain=(0.4, 0.3, 0.12), water=(0.12, 0.2, 0.12), l0=0.01,
                        wl=3, dl=2, dt=5e-4),
    "dam3d_mid": dict(dim=3, domain=(0.6, 0.4, 0.2), water=(0.2, 0.3, 0.2), l0=0.01,
                      wl=3, dl=2, dt=5e-4),
    # a small tank for the wall-friction fixture (7,300 particles, 320 fluid)
    "fric3d_tiny": dict(dim=3, domain=(0.2, 0.15, 0.05), water=(0.08, 0.1, 0.05), l0=0.01,
                        wl=3, dl=2, dt=5e-4),
}


def _py_round(q):
    """Python's round() to int (half to even) of a positive float."""
    return int(round(q))


def dambreak(dim, domain, water, l0, wl, dl, **_):
    """Returns (pos[n,dim], vel[n,dim], type[n]) in generator order."""
    layers = wl + dl
    if dim == 3:
        face = [int(domain[0] / l0), int(domain[1] / l0), int(domain[2] / l0)]
    else:
        face = [_py_round(domain[0] / l0), _py_round(domain[1] / l0)]
    lo = [-(layers - 1)] * dim
    hi = [face[d] + (layers - 1) if d != 1 else face[1] for d in range(dim)]
    axes = [np.arange(lo[d], hi[d] + 1) for d in range(dim)]
    grids = np.meshgrid(*axes, indexing="ij")
    idx = [g.ravel() for g in grids]
    t = np.full(idx[0].shape, -1, dtype=np.int32)
    for d in range(dim):
        t[idx[d] <= 0] = 1
        if d != 1:
            t[idx[d] >= face[d]] = 1
    for d in range(dim):
        t[idx[d] <= -wl] = 3
        if d != 1:
            t[idx[d] >= face[d] + wl] = 3
    eps = 1.0e-10
    fluid = t < 0
    for d in range(dim):
        fluid &= (idx[d] >= 1) & (idx[d] * l0 <= water[d] + eps)
    t[fluid] = 0
    keep = t >= 0
    pos = np.empty((int(keep.sum()), dim))
    for d in range(dim):
        vals = {}
        col = idx[d][keep]
        for v in np.unique(col):
            vals[int(v)] = float("%.8e" % (int(v) * l0))
        pos[:, d] = np.array([vals[int(v)] for v in col]) if len(col) else []
    return pos, np.zeros_like(pos), t[keep].astype(np.int32)


def scene_config(name):
    return dict(CONFIGS[name])


def params_for(name, **over):
    """Benchmark physics (influence ratios 3.1/2.1, density RHS, CG) for a
    named scene; keyword overrides win."""
    sc = CONFIGS[name]
    dim = sc["dim"]
    p = dict(particle_distance=sc["l0"], influence_ratio_lap=3.1, influence_ratio_n=2.1,
             max_neighbors=512, wall_layers=sc["wl"], dummy_layers=sc["dl"],
             density=1000.0, viscosity=1e-6, dt=sc["dt"], ppe_type=0, hs_mode=0,
             relaxation_coeff=0.2, surface_threshold=0.97, restitution_coeff=0.2,
             collision_distance_ratio=0.5, cmps_gradient=0, use_analytical_lambda=0,
             clamp_negative_pressure=1, solver_type=0, cg_tolerance=1e-9, cg_max_iter=20000,
             gravity=[0.0, -9.81, 0.0][:dim], domain_min=[0.0] * dim,
             domain_max=list(sc["domain"]))
    p.update(over)
    return p


def load_particles_txt(path, dim):
    """Reads a reference particles.txt (count line, then rows)."""
    rows = []
    count = None
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            if count is None:
                count = int(s.split()[0])
                continue
            rows.append(s.split())
            if len(rows) == count:
                break
    arr = np.array(rows, dtype=object)
    pos = arr[:, :dim].astype(float)
    vel = arr[:, dim:2 * dim].astype(float)
    typ = arr[:, 2 * dim].astype(np.int32)
    return pos, vel, typ
