"""Host side of libmps_hip without a GPU: the C-ABI loads and exports every
function include/*.h declares, and the host layer (cal/params parser,
particles loader, dam-break generator, CSV/VTK writers) behaves like the
reference's."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import scenes
from mps_amd import abi as A
from oracle.oracle import Reference, make_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(serial=False):
    """Function names include/mps_hip.h declares; those inside an
    #ifdef MPS_FRICTION_FIELDS block only for the 3-D serial layout."""
    names = set()
    for h in ("mps_hip.h",):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        if not serial:
            text = re.sub(r"#ifdef MPS_FRICTION_FIELDS.*?#endif", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, flags=re.M):
            if m.group(1) not in ("if", "defined"):
                names.add(m.group(1))
    return names


@pytest.fixture(scope="module", params=[(3, False), (2, False), (3, True)], ids=["3d", "2d", "3ds"])
def lib(request):
    import mps_amd
    return mps_amd.load(*request.param)


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions(serial=lib.abi.SimConfig is not None and hasattr(lib.abi.SimConfig, "friction_enabled"))
    assert len(names) > 50
    missing = [n for n in sorted(names) if not hasattr(lib.L, n)]
    assert not missing, missing
    assert lib.missing == []
    assert hasattr(lib.L, "g_config")


def test_no_device_reports_error_without_gpu(lib):
    import conftest
    if conftest.gpu_available():
        pytest.skip("a GPU is present")
    assert lib.mps_hip_init(0) == 2          # MPS_ERR_NO_DEVICE
    assert lib.mps_hip_device_arch() == b""


PARAMS_TEXT = """# every key of sim_config.c:166-270, plus noise
particle_distance    0.0125
influence_ratio_lap  3.3
influence_ratio_n    2.4
max_neighbors        300
wall_layers          3
dummy_layers         1
density              998.2
viscosity            1.1e-06
gravity_x            0.5
gravity_y            -9.80665
gravity_z            0.25
dt                   0.00025
t_end                0.75
output_interval      40
solver_type          1
cg_max_iter          777
cg_tolerance         2.5e-9
relaxation_coeff     0.3
clamp_negative_pressure 1
ppe_type             1
c_ppe                1.02
gamma_ppe            0.02
surface_threshold    0.95
surface_detection_method 1
surface_count_threshold  0.8
restitution_coeff        0.1
collision_distance_ratio 0.45
surface_tension_enabled  1
surface_tension_coeff    0.07
surface_tension_re_ratio 3.0
wetting_angle_SL         75.0
domain_x_min         -0.01
domain_x_max         1.01
domain_y_min         -0.02
domain_y_max         0.61
domain_z_min         -0.03
domain_z_max         0.52
cmps_gradient        2
use_analytical_lambda 1
hs_mode              1
output_dir           out_dir_x
not_a_key            5
   # indented comment

"""


def _load(lib_or_ref, abi, tmp_path, text=PARAMS_TEXT):
    cal = tmp_path / "case" / "cal.txt"
    cal.parent.mkdir(exist_ok=True)
    (cal.parent / "params.txt").write_text(text)
    cal.write_text("# cal\nparticle_file  particles.txt\nparam_file     params.txt\n")
    cfg = abi.SimConfig()
    L = lib_or_ref.L
    L.config_set_defaults(C.byref(cfg))
    assert L.config_load_cal(str(cal).encode(), C.byref(cfg)) == 0
    assert L.config_load_params(cfg.param_file, C.byref(cfg)) == 0
    return cfg


def _as_dict(cfg):
    out = {}
    for name, _ in cfg._fields_:
        v = getattr(cfg, name)
        out[name] = list(v) if hasattr(v, "_length_") and not isinstance(v, bytes) else v
    return out


@pytest.mark.skipif(not Reference.available("3dp"), reason="reference build absent")
@pytest.mark.parametrize("dim", [3, 2])
def test_config_parser_matches_reference(dim, tmp_path):
    import mps_amd
    lib = mps_amd.load(dim)
    ref = Reference("3dp" if dim == 3 else "2d")
    text = PARAMS_TEXT if dim == 3 else "\n".join(l for l in PARAMS_TEXT.splitlines() if "_z_" not in l and "gravity_z" not in l)
    mine = _as_dict(_load(lib, lib.abi, tmp_path, text))
    theirs = _as_dict(_load(ref, ref.abi, tmp_path, text))
    assert mine == theirs


@pytest.mark.skipif(not Reference.available("3dp"), reason="reference build absent")
def test_defaults_match_reference():
    import mps_amd
    for dim, v in ((3, "3dp"), (2, "2d")):
        lib = mps_amd.load(dim)
        ref = Reference(v)
        a, b = lib.abi.SimConfig(), ref.abi.SimConfig()
        lib.L.config_set_defaults(C.byref(a))
        ref.L.config_set_defaults(C.byref(b))
        assert _as_dict(a) == _as_dict(b)


def test_path_resolution(tmp_path):
    import mps_amd
    lib = mps_amd.load(3)
    cfg = _load(lib, lib.abi, tmp_path)
    assert cfg.particle_file.decode() == str(tmp_path / "case" / "particles.txt")
    assert cfg.influence_radius_lap == 3.3 * 0.0125
    assert cfg.influence_radius_n == 2.4 * 0.0125
    assert cfg.influence_radius_st == 3.0 * 0.0125


@pytest.mark.skipif(not Reference.available("3dp"), reason="reference build absent")
@pytest.mark.parametrize("dim", [3, 2])
def test_output_files_match_reference_bytes(dim, tmp_path):
    import mps_amd
    lib = mps_amd.load(dim)
    ref = Reference("3dp" if dim == 3 else "2d")
    name = "shipped3d" if dim == 3 else "shipped2d"
    pos, vel, typ = scenes.dambreak(**scenes.CONFIGS[name])
    rng = np.random.default_rng(0)
    s = A.HostScene(A.types(dim), pos, rng.normal(size=vel.shape), typ)
    s.arr["pressure"] = rng.normal(size=len(typ)) * 1e3
    s.arr["type"][rng.choice(len(typ), 40, replace=False)] = 2
    for who, d in ((lib.L, tmp_path / "mine"), (ref.L, tmp_path / "ref")):
        d.mkdir()
        who.output_csv.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
        who.output_vtk.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
        who.output_csv(C.cast(s.ptr(), C.c_void_p), 1200, str(d).encode())
        who.output_vtk(C.cast(s.ptr(), C.c_void_p), 1200, str(d).encode())
    for f in ("output_001200.csv", "output_001200.vtk"):
        assert (tmp_path / "mine" / f).read_bytes() == (tmp_path / "ref" / f).read_bytes(), f


@pytest.mark.parametrize("name", ["shipped3d", "dam3d_small", "shipped2d", "dam2d_l0.01", "dam2d_l0.005"])
def test_c_generator_matches_lattice_rules(name, tmp_path):
    import mps_amd
    sc = scenes.CONFIGS[name]
    dim = sc["dim"]
    lib = mps_amd.load(dim)
    d = list(sc["domain"]) + ([0.0] if dim == 2 else [])
    w = list(sc["water"]) + ([0.0] if dim == 2 else [])
    dims = (C.c_double * 6)(*(d[:3] + w[:3]))
    path = tmp_path / "particles.txt"
    n = lib.mps_gen_dambreak(dims, sc["l0"], sc["wl"], sc["dl"], None, str(path).encode())
    ps = lib.particle_system_create(n)
    assert lib.mps_gen_dambreak(dims, sc["l0"], sc["wl"], sc["dl"], ps, None) == n
    arr = np.ctypeslib.as_array(C.cast(ps.contents.particles, C.POINTER(lib.abi.Particle)), shape=(n,))
    pos, vel, typ = scenes.dambreak(**sc)
    assert np.array_equal(arr["pos"], pos) and np.array_equal(arr["type"], typ)
    # the text file round-trips through the loader
    assert lib.mps_count_particles(str(path).encode()) == n
    ps2 = lib.particle_system_create(n)
    assert lib.mps_load_particles(str(path).encode(), ps2) == 0
    arr2 = np.ctypeslib.as_array(C.cast(ps2.contents.particles, C.POINTER(lib.abi.Particle)), shape=(n,))
    assert np.array_equal(arr2["pos"], pos) and np.array_equal(arr2["type"], typ)
    lib.particle_system_free(ps)
    lib.particle_system_free(ps2)


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="reference inputs absent")
def test_generator_reproduces_shipped_inputs():
    for name, path, dim in (("shipped3d", "mps_fluid3D_parallel/examples/dambreak/particles.txt", 3),
                            ("shipped2d", "mps_fluid2D/examples/dambreak/particles.txt", 2)):
        pos, vel, typ = scenes.dambreak(**scenes.CONFIGS[name])
        P, V, T = scenes.load_particles_txt(os.path.join("/root/reference", path), dim)
        assert np.array_equal(pos, P) and np.array_equal(typ, T)


def test_kernel_weight_host(lib):
    cfg = make_config(lib.abi, particle_distance=0.01)
    lib.set_config(cfg)
    re = 0.031
    for r in (0.0, 1e-5, 1e-4, 0.005, 0.0309999, 0.031, 0.05):
        want = 0.0 if r >= re else re / max(r, 0.01 * 0.01) - 1.0
        assert lib.kernel_weight(r, re) == want


@pytest.mark.parametrize("dim", [3, 2])
def test_driver_sees_its_config(dim, tmp_path):
    """The stand-alone driver (the reference main.c's flow, linked against
    the library) sets g_config in the executable; the library must read
    that same variable, so simulation_run starts (it prints the device line
    first, before any GPU work) instead of returning for want of a
    config.  Runs without a GPU: the run stops at the device check."""
    import subprocess
    import mps_amd
    sc = scenes.CONFIGS["dam3d_small" if dim == 3 else "dam2d_l0.01"]
    lib = mps_amd.load(dim)
    d = list(sc["domain"]) + ([0.0] if dim == 2 else [])
    w = list(sc["water"]) + ([0.0] if dim == 2 else [])
    dims = (C.c_double * 6)(*(d[:3] + w[:3]))
    assert lib.mps_gen_dambreak(dims, sc["l0"], sc["wl"], sc["dl"], None,
                                str(tmp_path / "particles.txt").encode()) > 0
    (tmp_path / "params.txt").write_text("particle_distance %r\ndt 5e-4\nt_end 1e-3\noutput_interval 1\n"
                                         "output_dir out\n" % sc["l0"])
    (tmp_path / "cal.txt").write_text("particle_file particles.txt\nparam_file params.txt\n")
    driver = os.path.join(os.path.dirname(mps_amd.lib_path(dim)), "mps_sim%dd" % dim)
    p = subprocess.run([driver, "cal.txt"], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert "Loaded particles" in p.stdout
    assert "HIP device:" in p.stdout, (p.stdout[-800:], p.stderr[-800:])


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="reference sources absent")
@pytest.mark.parametrize("prog,libname", [("mps_fluid3D_parallel", "libmps_hip3d.so"),
                                          ("mps_fluid2D", "libmps_hip2d.so"),
                                          ("mps_fluid3D", "libmps_hip3ds.so")])
def test_reference_main_relinks(prog, libname, tmp_path):
    """Each reference program's own main.c, compiled here and linked
    against the matching library as the integration notes describe, loads
    its shipped dam break and enters simulation_run with its configuration
    (without a GPU the run stops at the device check)."""
    import shutil
    import subprocess
    import mps_amd
    src = os.path.join("/root/reference", prog)
    libdir = os.path.dirname(mps_amd.lib_path(3))
    obj, exe = tmp_path / "main.o", tmp_path / "main_hip"
    subprocess.run(["gcc", "-O2", "-w", "-I" + os.path.join(src, "include"), "-c",
                    os.path.join(src, "src", "main.c"), "-o", str(obj)], check=True)
    subprocess.run(["gcc", "-o", str(exe), str(obj), "-L" + libdir, "-l:" + libname,
                    "-Wl,-rpath," + libdir, "-lm"], check=True)
    case = tmp_path / "case"
    case.mkdir()
    for f in ("cal.txt", "params.txt", "particles.txt"):
        shutil.copy(os.path.join(src, "examples", "dambreak", f), case / f)
    # two steps, so that a host with a GPU runs a short case too
    import re
    par = (case / "params.txt").read_text()
    dt = float(re.search(r"^dt\s+(\S+)", par, re.M).group(1))
    (case / "params.txt").write_text(re.sub(r"^t_end\s+\S+", "t_end %r" % (2.5 * dt), par, flags=re.M))
    p = subprocess.run([str(exe), "cal.txt"], cwd=str(case), capture_output=True, text=True, timeout=300)
    assert "Loaded particles" in p.stdout
    assert "HIP device:" in p.stdout, (p.stdout[-800:], p.stderr[-800:])
