"""ctypes wrapper of the CPU oracle (liboracle{2,3}d.so) and of the reference
compiled from its sources (oracle/_ref/libref*.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker -- never by the product package.
"""
import ctypes as C
import os

import numpy as np

from mps_amd import abi as _abi

HERE = os.path.dirname(os.path.abspath(__file__))

DEFAULTS = dict(
    # config_set_defaults, mps_fluid3D_parallel/src/sim_config.c:20-79
    particle_distance=0.025, influence_ratio_lap=2.1, influence_ratio_n=2.1,
    max_neighbors=512, wall_layers=2, dummy_layers=2, density=1000.0,
    viscosity=1.0e-6, dt=5.0e-4, t_end=2.0, output_interval=100, solver_type=0,
    cg_max_iter=10000, cg_tolerance=1.0e-8, relaxation_coeff=0.2,
    clamp_negative_pressure=0, ppe_type=0, c_ppe=1.01, gamma_ppe=0.01,
    surface_threshold=0.97, surface_detection_method=0, surface_count_threshold=0.85,
    restitution_coeff=0.2, collision_distance_ratio=0.5, surface_tension_enabled=0,
    surface_tension_coeff=0.0728, surface_tension_re_ratio=3.2, wetting_angle_SL=90.0,
    cmps_gradient=0, use_analytical_lambda=0, hs_mode=0,
)


def make_config(abi, gravity=None, domain_min=None, domain_max=None, **kw):
    """SimConfig with the reference defaults, overrides, and the derived radii
    of config_load_params (sim_config.c:272-274)."""
    c = abi.SimConfig()
    vals = dict(DEFAULTS)
    vals.update(kw)
    for k, v in vals.items():
        setattr(c, k, v)
    dim = abi.dim
    g = gravity if gravity is not None else ([0.0, -9.81, 0.0][:dim])
    for d in range(dim):
        c.gravity[d] = g[d]
    dmin = domain_min if domain_min is not None else [0.0] * dim
    dmax = domain_max if domain_max is not None else [0.4, 0.3, 0.2][:dim]
    for d in range(dim):
        c.domain_min[d] = dmin[d]
        c.domain_max[d] = dmax[d]
    c.influence_radius_lap = c.influence_ratio_lap * c.particle_distance
    c.influence_radius_n = c.influence_ratio_n * c.particle_distance
    c.influence_radius_st = c.surface_tension_re_ratio * c.particle_distance
    c.output_dir = b"output"
    return c


class Oracle:
    """The C restatement (liboracle{dim}d.so; liboracle3ds.so in the 3-D
    serial program's layout, with its hybrid wall friction)."""

    def __init__(self, dim=3, friction=False):
        self.dim = dim
        self.abi = _abi.types(dim, friction)
        path = os.path.join(HERE, "liboracle%dd%s.so" % (dim, "s" if friction else ""))
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: %s (run make -C oracle)" % path)
        L = C.CDLL(path)
        self.L = L
        PS = C.POINTER(self.abi.ParticleSystem)
        CF = C.POINTER(self.abi.SimConfig)
        vp = C.c_void_p
        L.orc_sim_create.restype = vp
        L.orc_sim_free.argtypes = [vp]
        L.orc_sim_begin.argtypes = [vp, PS, CF]
        L.orc_step.argtypes = [vp, PS, CF]
        L.orc_init_params.argtypes = [PS, CF]
        L.orc_sim_set_opts.argtypes = [vp, C.c_int, C.c_int]
        L.orc_sim_nlist.argtypes = [vp, C.POINTER(C.POINTER(C.c_long)), C.POINTER(C.POINTER(C.c_int))]
        L.orc_sim_ppe.argtypes = [vp]
        L.orc_sim_ppe.restype = vp
        L.orc_weight.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_weight.restype = C.c_double
        L.orc_solve_csr.argtypes = [C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_int),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int,
                                    C.c_double, vp, C.POINTER(C.c_double)]
        L.orc_solve_pressure.argtypes = [vp, PS, CF]
        L.orc_load_particles.argtypes = [C.c_char_p, PS]
        if friction:
            L.orc_apply_friction.argtypes = [PS, CF]

    def sim(self, tol_mode=0, dot_order=0):
        return OracleSim(self, tol_mode, dot_order)

    def init_params(self, scene, cfg):
        self.L.orc_init_params(scene.ptr(), C.byref(cfg))

    def apply_friction(self, scene, cfg):
        """The hybrid wall friction alone (serial layout), on the scene's
        velocities over the list of its positions."""
        if self.L.orc_apply_friction(scene.ptr(), C.byref(cfg)) != 0:
            raise RuntimeError("orc_apply_friction failed")

    def solve_csr(self, row_ptr, col, val, b, solver=0, dense_ic=None, max_iter=10000,
                  tol=1e-10, tol_mode=0, dot_order=0):
        n = len(row_ptr) - 1
        rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        cc = np.ascontiguousarray(col, dtype=np.int32)
        vv = np.ascontiguousarray(val, dtype=np.float64)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(n)
        res = C.c_double()

        class Opts(C.Structure):
            _fields_ = [("tol_mode", C.c_int), ("dot_order", C.c_int)]
        o = Opts(tol_mode, dot_order)
        if dense_ic is None:
            dense_ic = 1 if self.dim == 2 else 0
        it = self.L.orc_solve_csr(n, rp.ctypes.data_as(C.POINTER(C.c_long)),
                                  cc.ctypes.data_as(C.POINTER(C.c_int)),
                                  vv.ctypes.data_as(C.POINTER(C.c_double)),
                                  bb.ctypes.data_as(C.POINTER(C.c_double)),
                                  x.ctypes.data_as(C.POINTER(C.c_double)),
                                  solver, dense_ic, max_iter, tol, C.cast(C.byref(o), C.c_void_p),
                                  C.byref(res))
        return x, it, res.value


class _PPE(C.Structure):
    _fields_ = [("n_eq", C.c_int), ("nnz", C.c_long), ("eq_idx", C.POINTER(C.c_int)),
                ("row_ptr", C.POINTER(C.c_long)), ("col", C.POINTER(C.c_int)),
                ("val", C.POINTER(C.c_double)), ("diag", C.POINTER(C.c_double)),
                ("b", C.POINTER(C.c_double)), ("x", C.POINTER(C.c_double)),
                ("l_ptr", C.POINTER(C.c_long)), ("l_col", C.POINTER(C.c_int)),
                ("l_val", C.POINTER(C.c_double)), ("l_diag", C.POINTER(C.c_double)),
                ("iters", C.c_int), ("res", C.c_double), ("bnorm", C.c_double),
                ("breakdown", C.c_int)]


class OracleSim:
    def __init__(self, orc, tol_mode=0, dot_order=0):
        self.orc = orc
        self.h = orc.L.orc_sim_create()
        orc.L.orc_sim_set_opts(self.h, tol_mode, dot_order)

    def __del__(self):
        try:
            self.orc.L.orc_sim_free(self.h)
        except Exception:
            pass

    def begin(self, scene, cfg):
        return self.orc.L.orc_sim_begin(self.h, scene.ptr(), C.byref(cfg))

    def step(self, scene, cfg):
        return self.orc.L.orc_step(self.h, scene.ptr(), C.byref(cfg))

    def solve_pressure(self, scene, cfg):
        return self.orc.L.orc_solve_pressure(self.h, scene.ptr(), C.byref(cfg))

    def neighbor_lists(self):
        st = C.POINTER(C.c_long)()
        ix = C.POINTER(C.c_int)()
        n = self.orc.L.orc_sim_nlist(self.h, C.byref(st), C.byref(ix))
        start = np.ctypeslib.as_array(st, shape=(n + 1,)).copy()
        idx = np.ctypeslib.as_array(ix, shape=(max(int(start[-1]), 1),))[:start[-1]].copy()
        return start, idx

    def ppe(self, num):
        q = _PPE.from_address(self.orc.L.orc_sim_ppe(self.h))
        out = {"n_eq": q.n_eq, "nnz": q.nnz, "iters": q.iters, "res": q.res,
               "bnorm": q.bnorm, "breakdown": q.breakdown}
        ne = q.n_eq
        out["eq_idx"] = np.ctypeslib.as_array(q.eq_idx, shape=(num,)).copy() if num else np.zeros(0, np.int32)
        if ne > 0:
            out["row_ptr"] = np.ctypeslib.as_array(q.row_ptr, shape=(ne + 1,)).copy()
            out["col"] = np.ctypeslib.as_array(q.col, shape=(q.nnz,)).copy()
            out["val"] = np.ctypeslib.as_array(q.val, shape=(q.nnz,)).copy()
            out["diag"] = np.ctypeslib.as_array(q.diag, shape=(ne,)).copy()
            out["b"] = np.ctypeslib.as_array(q.b, shape=(ne,)).copy()
            out["x"] = np.ctypeslib.as_array(q.x, shape=(ne,)).copy()
            if q.l_ptr:
                lp = np.ctypeslib.as_array(q.l_ptr, shape=(ne + 1,)).copy()
                out["l_ptr"] = lp
                m = int(lp[-1])
                out["l_col"] = np.ctypeslib.as_array(q.l_col, shape=(max(m, 1),))[:m].copy()
                out["l_val"] = np.ctypeslib.as_array(q.l_val, shape=(max(m, 1),))[:m].copy()
                out["l_diag"] = np.ctypeslib.as_array(q.l_diag, shape=(ne,)).copy()
        return out


class Reference:
    """The reference built from its own sources (oracle/_ref).  Available only
    where /root/reference exists and `make -C oracle ref` ran."""

    VARIANTS = {"3dp": ("libref3dp.so", 3, False), "3d": ("libref3d.so", 3, True),
                "2d": ("libref2d.so", 2, False)}

    @classmethod
    def available(cls, variant="3dp"):
        return os.path.exists(os.path.join(HERE, "_ref", cls.VARIANTS[variant][0]))

    def __init__(self, variant="3dp"):
        name, dim, friction = self.VARIANTS[variant]
        self.dim = dim
        self.abi = _abi.types(dim, friction)
        L = C.CDLL(os.path.join(HERE, "_ref", name))
        self.L = L
        a = self.abi
        PS = C.POINTER(a.ParticleSystem)
        NL = C.POINTER(a.NeighborList)
        CL = C.POINTER(a.CellList)
        L.neighbor_list_create.argtypes = [C.c_int, C.c_int]
        L.neighbor_list_create.restype = NL
        L.neighbor_list_free.argtypes = [NL]
        L.cell_list_create.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.cell_list_create.restype = CL
        L.cell_list_free.argtypes = [CL]
        L.neighbor_search_cell_linked_list.argtypes = [NL, PS, CL, C.c_double]
        L.neighbor_search_brute_force.argtypes = [NL, PS, C.c_double]
        for f in ("calc_particle_number_density", "calc_viscosity_term",
                  "calc_pressure_gradient", "solve_pressure", "calc_neighbor_count"):
            getattr(L, f).argtypes = [PS, NL]
        if dim == 3:
            L.collision.argtypes = [PS, NL]
        else:
            L.collision.argtypes = [PS]
        L.judge_free_surface.argtypes = [PS, C.c_double]
        L.clamp_pressure.argtypes = [PS]
        L.apply_wall_boundary.argtypes = [PS]
        L.particle_system_calc_initial_params.argtypes = [PS]
        L.simulation_step.argtypes = [PS, NL, CL, C.c_int]
        L.kernel_weight.argtypes = [C.c_double, C.c_double]
        L.kernel_weight.restype = C.c_double
        self._g = C.c_void_p.in_dll(L, "g_config")
        self._cfg = None

    def set_config(self, cfg):
        self._cfg = cfg
        self._g.value = C.addressof(cfg)

    def config_from(self, src_cfg):
        """Copies a SimConfig of another layout field by field."""
        c = self.abi.SimConfig()
        for name, _ in self.abi.SimConfig._fields_:
            if hasattr(src_cfg, name):
                v = getattr(src_cfg, name)
                if hasattr(v, "_length_"):
                    arr = getattr(c, name)
                    for k in range(len(arr)):
                        arr[k] = v[k]
                else:
                    setattr(c, name, v)
        return c

    def search_radius(self):
        c = self._cfg
        re = max(c.influence_radius_lap, c.influence_radius_n)
        if c.surface_tension_enabled:
            re = max(re, c.influence_radius_st)
        return re

    def make_lists(self, n):
        c = self._cfg
        re = self.search_radius()
        dmin = (C.c_double * self.dim)(*[c.domain_min[d] for d in range(self.dim)])
        dmax = (C.c_double * self.dim)(*[c.domain_max[d] for d in range(self.dim)])
        nl = self.L.neighbor_list_create(n, c.max_neighbors)
        cl = self.L.cell_list_create(n, re, dmin, dmax)
        return nl, cl

    def free_lists(self, nl, cl):
        self.L.neighbor_list_free(nl)
        self.L.cell_list_free(cl)
