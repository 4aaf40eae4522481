"""mps_amd -- ctypes binding of libmps_hip{2,3}d.so, the MI355X (gfx950)
implementation of the MPS timestep behind the reference's C entry points.

The library is the product; this module only loads it and declares its
prototypes (include/mps_hip.h).  There is no Python or CPU fallback: if the
library is missing, loading raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))


def lib_path(dim=3, serial=False):
    """libmps_hip3d.so / libmps_hip2d.so, or libmps_hip3ds.so (the 3-D serial
    program's SimConfig layout, with its hybrid wall friction)."""
    return os.path.join(HERE, "libmps_hip%dd%s.so" % (dim, "s" if serial else ""))


class StepStats(C.Structure):
    _fields_ = [("steps", C.c_int), ("last_iters", C.c_int), ("total_iters", C.c_longlong),
                ("last_outer", C.c_int), ("n_eq", C.c_int), ("nnz", C.c_longlong),
                ("n_pairs", C.c_longlong), ("last_residual", C.c_double),
                ("last_bnorm", C.c_double), ("ic_levels", C.c_int), ("max_count", C.c_int)]


class TuneResult(C.Structure):
    _fields_ = [("spmv", C.c_int), ("check_every", C.c_int), ("spmv_ms", C.c_double * 3),
                ("solve_ms", C.c_double * 4), ("iters", C.c_int), ("pblk", C.c_int),
                ("pblk_ms", C.c_double * 3), ("search", C.c_int), ("search_ms", C.c_double * 2)]


_libs = {}


class Lib:
    """Prototyped handle on libmps_hip{dim}d.so."""

    def __init__(self, dim=3, serial=False):
        path = lib_path(dim, serial)
        if not os.path.exists(path):
            raise RuntimeError("%s is not built (python -m mps_amd.build)" % path)
        self.dim = dim
        self.abi = abi.types(dim, serial)
        self.L = C.CDLL(path)
        a = self.abi
        L = self.L
        PS = C.POINTER(a.ParticleSystem)
        NL = C.POINTER(a.NeighborList)
        CL = C.POINTER(a.CellList)
        CF = C.POINTER(a.SimConfig)
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int)
        sig = {
            "config_set_defaults": (None, [CF]),
            "config_load_cal": (C.c_int, [C.c_char_p, CF]),
            "config_load_params": (C.c_int, [C.c_char_p, CF]),
            "config_print": (None, [CF]),
            "particle_system_create": (PS, [C.c_int]),
            "particle_system_free": (None, [PS]),
            "particle_system_add": (C.c_int, [PS, D, D, C.c_int]),
            "particle_system_calc_initial_params": (None, [PS]),
            "kernel_weight": (C.c_double, [C.c_double, C.c_double]),
            "kernel_weight_derivative": (C.c_double, [C.c_double, C.c_double]),
            "neighbor_list_create": (NL, [C.c_int, C.c_int]),
            "neighbor_list_free": (None, [NL]),
            "cell_list_create": (CL, [C.c_int, C.c_double, D, D]),
            "cell_list_free": (None, [CL]),
            "neighbor_search_brute_force": (None, [NL, PS, C.c_double]),
            "neighbor_search_cell_linked_list": (None, [NL, PS, CL, C.c_double]),
            "neighbor_get": (C.c_int, [NL, C.c_int, C.c_int]),
            "calc_particle_number_density": (None, [PS, NL]),
            "calc_viscosity_term": (None, [PS, NL]),
            "calc_pressure_gradient": (None, [PS, NL]),
            "judge_free_surface": (None, [PS, C.c_double]),
            "calc_neighbor_count": (None, [PS, NL]),
            "judge_free_surface_by_count": (None, [PS, C.c_double]),
            "clamp_pressure": (None, [PS]),
            "collision": (None, [PS, NL] if dim == 3 else [PS]),
            "calc_surface_tension": (None, [PS, NL]),
            "calc_surface_tension_SL": (None, [PS, NL]),
            "apply_wall_boundary": (None, [PS]),
            "remove_out_of_bounds": (None, [PS] + [C.c_double] * (2 * dim)),
            "solve_pressure": (None, [PS, NL]),
            "simulation_run": (None, [PS]),
            "simulation_step": (None, [PS, NL, CL, C.c_int]),
            "output_csv": (None, [PS, C.c_int, C.c_char_p]),
            "output_vtk": (None, [PS, C.c_int, C.c_char_p]),
            "mps_load_particles": (C.c_int, [C.c_char_p, PS]),
            "mps_count_particles": (C.c_int, [C.c_char_p]),
            "mps_gen_dambreak": (C.c_int, [D, C.c_double, C.c_int, C.c_int, PS, C.c_char_p]),
            "mps_hip_init": (C.c_int, [C.c_int]),
            "mps_hip_finalize": (None, []),
            "mps_hip_last_error": (C.c_int, []),
            "mps_hip_error_string": (C.c_char_p, [C.c_int]),
            "mps_hip_device_arch": (C.c_char_p, []),
            "mps_hip_upload": (C.c_int, [PS, NL]),
            "mps_hip_download": (C.c_int, [PS, NL]),
            "mps_hip_set_resident": (C.c_int, [PS, C.c_int]),
            "mps_hip_release": (None, [PS]),
            "mps_hip_run_steps": (C.c_int, [PS, C.c_int]),
            "mps_hip_run_batch": (C.c_int, [C.POINTER(PS), C.POINTER(C.POINTER(a.SimConfig)), C.c_int, C.c_int]),
            "mps_hip_adjoint": (C.c_int, [PS, C.c_int, C.c_int, D, D, D, D, D]),
            "mps_hip_replay_losses": (C.c_int, [PS, C.c_int, D, C.c_int, D, D, D, D]),
            "mps_hip_step_backward": (C.c_int, [PS, NL, D, D, D, D, D, D]),
            "mps_hip_set_option": (C.c_int, [C.c_char_p, C.c_double]),
            "mps_hip_get_option": (C.c_double, [C.c_char_p]),
            "mps_hip_get_stats": (C.c_int, [PS, C.POINTER(StepStats)]),
            "mps_hip_autotune": (C.c_int, [PS, C.c_int, C.POINTER(TuneResult)]),
            "mps_hip_autotune_batch": (C.c_int, [C.POINTER(PS), C.POINTER(C.POINTER(a.SimConfig)), C.c_int, C.c_int,
                                                 C.POINTER(TuneResult)]),
            "mps_hip_export_ppe": (C.c_int, [PS, I, I, I, D, D, D]),
            "mps_hip_export_solution": (C.c_int, [PS, D]),
            "mps_hip_export_ic": (C.c_int, [PS, I, I, D, D]),
            "mps_hip_profile_get": (C.c_int, [C.c_char_p, D, C.POINTER(C.c_longlong)]),
            "mps_hip_profile_reset": (None, []),
            "mps_hip_solve_csr": (C.c_int, [C.c_int, I, I, D, D, D, C.c_int, C.c_int, C.c_double, I]),
            "mps_hip_solve_csr_noalloc": (C.c_int, [C.c_int, I, I, D, D, D, C.c_int, C.c_int, C.c_double, I]),
            "mps_hip_fp64_probe": (C.c_int, [C.c_int, D, D, D, D]),
        }
        self.missing = []
        if serial:
            sig["apply_friction"] = (None, [PS, NL])
        for name, (res, args) in sig.items():
            try:
                f = getattr(L, name)
            except AttributeError:
                self.missing.append(name)
                continue
            f.restype = res
            f.argtypes = args
        self._g = C.c_void_p.in_dll(L, "g_config")
        self._cfg = None

    def __getattr__(self, name):
        return getattr(self.L, name)

    def set_config(self, cfg):
        """Points the library's g_config at cfg (kept alive here)."""
        self._cfg = cfg
        self._g.value = C.addressof(cfg)

    def set_option(self, key, value):
        rc = self.L.mps_hip_set_option(key.encode(), float(value))
        if rc:
            raise KeyError(key)

    def check(self, rc, what="call"):
        if rc:
            msg = self.L.mps_hip_error_string(rc).decode()
            raise RuntimeError("%s failed: %s (%d)" % (what, msg, rc))

    def stats(self, scene):
        """Step statistics of a HostScene or a ParticleSystem pointer."""
        st = StepStats()
        ps = scene.ptr() if hasattr(scene, "ptr") else scene
        self.check(self.L.mps_hip_get_stats(ps, C.byref(st)), "get_stats")
        return st

    def export_ppe(self, scene):
        st = self.stats(scene)
        n = scene.num
        ne = st.n_eq
        nnz = st.nnz
        eq = np.zeros(n, np.int32)
        rp = np.zeros(ne + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1), np.float64)
        diag = np.zeros(max(ne, 1), np.float64)
        b = np.zeros(max(ne, 1), np.float64)
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        self.check(self.L.mps_hip_export_ppe(scene.ptr(), P(eq, C.c_int), P(rp, C.c_int),
                                             P(col, C.c_int), P(val, C.c_double),
                                             P(diag, C.c_double), P(b, C.c_double)), "export_ppe")
        return dict(eq_idx=eq, row_ptr=rp, col=col[:nnz], val=val[:nnz], diag=diag[:ne],
                    b=b[:ne], n_eq=ne, nnz=nnz)

    def export_solution(self, scene, n_eq):
        x = np.zeros(max(n_eq, 1))
        self.check(self.L.mps_hip_export_solution(scene.ptr(), x.ctypes.data_as(C.POINTER(C.c_double))),
                   "export_solution")
        return x[:n_eq]


def load(dim=3, serial=False):
    key = (dim, serial)
    if key not in _libs:
        _libs[key] = Lib(dim, serial)
    return _libs[key]
