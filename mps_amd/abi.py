"""ctypes mirrors of the C types in include/mps_types.h.

The layouts are those of the reference programs (mps_fluid3D_parallel/include/
particle.h:7-28, neighbor_search.h:8-30, sim_config.h:10-80; 2-D:
mps_fluid2D/include/*.h; with friction fields: mps_fluid3D/include/
sim_config.h:72-77, after hs_mode), so one Python-side ParticleSystem can be handed to this
library and to the reference compiled from its sources alike.
"""
import ctypes as C

import numpy as np

FLUID, WALL, GHOST, DUMMY = 0, 1, 2, 3

_cache = {}


def types(dim=3, friction=False):
    """Returns a namespace of ctypes classes for the given layout."""
    key = (dim, friction)
    if key in _cache:
        return _cache[key]
    if friction:
        # only SimConfig differs; the particle/list types are shared
        base = types(dim, False)
        Particle, ParticleSystem = base.Particle, base.ParticleSystem
        NeighborList, CellList = base.NeighborList, base.CellList
    else:
        Particle = ParticleSystem = NeighborList = CellList = None
    if Particle is None:
        Particle, ParticleSystem, NeighborList, CellList = _particle_types(dim)

    cfg = _config_fields(dim, friction)

    class SimConfig(C.Structure):
        _fields_ = cfg

    ns = type("ABI", (), {})()
    ns.dim = dim
    ns.Particle = Particle
    ns.ParticleSystem = ParticleSystem
    ns.NeighborList = NeighborList
    ns.CellList = CellList
    ns.SimConfig = SimConfig
    _cache[key] = ns
    return ns


def _particle_types(dim):
    class Particle(C.Structure):
        _fields_ = [("pos", C.c_double * dim), ("vel", C.c_double * dim),
                    ("acc", C.c_double * dim), ("pressure", C.c_double),
                    ("n", C.c_double), ("neighbor_count", C.c_int),
                    ("type", C.c_int), ("on_surface", C.c_int)]

    class ParticleSystem(C.Structure):
        _fields_ = [("particles", C.POINTER(Particle)), ("num", C.c_int),
                    ("capacity", C.c_int), ("n0", C.c_double),
                    ("lambda_", C.c_double), ("C_LL", C.c_double),
                    ("C_SL", C.c_double), ("n0_count", C.c_int)]

    class NeighborList(C.Structure):
        _fields_ = [("neighbors", C.POINTER(C.c_int)), ("count", C.POINTER(C.c_int)),
                    ("max_neighbors", C.c_int)]

    cell_fields = [("head", C.POINTER(C.c_int)), ("next", C.POINTER(C.c_int)),
                   ("nx", C.c_int), ("ny", C.c_int)]
    if dim == 3:
        cell_fields.append(("nz", C.c_int))
    cell_fields += [("total_cells", C.c_int), ("cell_size", C.c_double),
                    ("origin", C.c_double * dim)]

    class CellList(C.Structure):
        _fields_ = cell_fields

    return Particle, ParticleSystem, NeighborList, CellList


def _config_fields(dim, friction):
    cfg = [("particle_distance", C.c_double), ("influence_ratio_lap", C.c_double),
           ("influence_radius_lap", C.c_double), ("influence_ratio_n", C.c_double),
           ("influence_radius_n", C.c_double), ("max_neighbors", C.c_int),
           ("wall_layers", C.c_int), ("dummy_layers", C.c_int),
           ("density", C.c_double), ("viscosity", C.c_double),
           ("gravity", C.c_double * dim), ("dt", C.c_double), ("t_end", C.c_double),
           ("output_interval", C.c_int), ("solver_type", C.c_int),
           ("cg_max_iter", C.c_int), ("cg_tolerance", C.c_double),
           ("relaxation_coeff", C.c_double), ("clamp_negative_pressure", C.c_int),
           ("ppe_type", C.c_int), ("c_ppe", C.c_double), ("gamma_ppe", C.c_double),
           ("surface_threshold", C.c_double), ("surface_detection_method", C.c_int),
           ("surface_count_threshold", C.c_double), ("restitution_coeff", C.c_double),
           ("collision_distance_ratio", C.c_double), ("surface_tension_enabled", C.c_int),
           ("surface_tension_coeff", C.c_double), ("surface_tension_re_ratio", C.c_double),
           ("influence_radius_st", C.c_double), ("wetting_angle_SL", C.c_double)]
    cfg += [("domain_min", C.c_double * dim), ("domain_max", C.c_double * dim),
            ("cmps_gradient", C.c_int), ("use_analytical_lambda", C.c_int),
            ("hs_mode", C.c_int)]
    if friction:
        cfg += [("friction_enabled", C.c_int), ("friction_delta_theta", C.c_double),
                ("dynamic_friction_alpha", C.c_double),
                ("friction_velocity_threshold", C.c_double)]
    cfg += [("output_dir", C.c_char * 256),
            ("particle_file", C.c_char * 256), ("param_file", C.c_char * 256)]
    return cfg


class HostScene:
    """A ParticleSystem whose particle array is owned by Python.

    `arr` is a structured numpy view of the ctypes Particle array, so
    positions etc. can be read and written as numpy arrays."""

    def __init__(self, abi, pos, vel, ptype, capacity=None):
        n = len(ptype)
        cap = capacity or n
        self.abi = abi
        self.buf = (abi.Particle * max(cap, 1))()
        self.ps = abi.ParticleSystem()
        self.ps.particles = C.cast(self.buf, C.POINTER(abi.Particle))
        self.ps.num = n
        self.ps.capacity = cap
        self.arr = np.ctypeslib.as_array(self.buf)[:n]
        if n:
            self.arr["pos"][:] = pos
            self.arr["vel"][:] = vel
            self.arr["type"][:] = ptype

    @property
    def num(self):
        return self.ps.num

    def ptr(self):
        return C.byref(self.ps)

    def copy(self):
        c = HostScene(self.abi, self.arr["pos"], self.arr["vel"], self.arr["type"])
        c.arr[:] = self.arr
        for f in ("n0", "lambda_", "C_LL", "C_SL", "n0_count"):
            setattr(c.ps, f, getattr(self.ps, f))
        return c

    def snapshot(self):
        return self.arr.copy()


def padded_lists(nl, n):
    """Python lists of the neighbour rows of a padded NeighborList."""
    cnt = np.ctypeslib.as_array(nl.count, shape=(n,)).copy()
    mx = nl.max_neighbors
    flat = np.ctypeslib.as_array(nl.neighbors, shape=(n * mx,))
    return [flat[i * mx:i * mx + cnt[i]].copy() for i in range(n)]
