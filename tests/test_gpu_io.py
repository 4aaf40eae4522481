"""simulation_run's snapshots (the reference's CSV and legacy VTK formats,
mps_fluid3D_parallel/src/simulation.c:123-181, io.c:10-107): the writer that
formats a pinned host copy on its own thread while the device steps on
(option async_output 1, the default) writes the same files, byte for byte,
as the synchronous writer, and leaves the same final state."""
import os
import sys

import numpy as np
import pytest

import harness as H
from mps_amd import abi as A

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(lib, name, out_dir, async_output, steps, every):
    dim, cfg, pos, vel, typ = H.build(name, cg_tolerance=1e-10)
    cfg.t_end = (steps + 0.5) * cfg.dt          # (int)(t_end / dt) == steps
    cfg.output_interval = every
    cfg.output_dir = out_dir.encode()
    os.makedirs(out_dir, exist_ok=True)
    lib.set_config(cfg)
    lib.set_option("tol_mode", 1)
    lib.set_option("async_output", async_output)
    sc = A.HostScene(lib.abi, pos, vel, typ)
    lib.particle_system_calc_initial_params(sc.ptr())
    say("simulation_run", name, "async_output", async_output, "steps", steps, "every", every)
    lib.simulation_run(sc.ptr())
    assert lib.mps_hip_last_error() == 0
    lib.mps_hip_release(sc.ptr())
    return sc


@pytest.mark.parametrize("libname,name", [("lib3", "dam3d_small"), ("lib2", "dam2d_l0.01")])
def test_async_snapshots_same_bytes(request, tmp_path, libname, name):
    lib = request.getfixturevalue(libname)
    try:
        a = _run(lib, name, str(tmp_path / "sync"), 0, steps=7, every=2)
        b = _run(lib, name, str(tmp_path / "async"), 1, steps=7, every=2)
    finally:
        lib.set_option("async_output", 1)
        lib.set_option("tol_mode", 0)
    fa = sorted(os.listdir(tmp_path / "sync"))
    fb = sorted(os.listdir(tmp_path / "async"))
    want = ["output_%06d.%s" % (s, e) for s in (0, 2, 4, 6) for e in ("csv", "vtk")]
    assert fa == sorted(want)
    assert fb == fa
    for f in fa:
        x = (tmp_path / "sync" / f).read_bytes()
        y = (tmp_path / "async" / f).read_bytes()
        assert len(x) > 0 and x == y, f
    # the snapshots are of different states, and the last one differs from
    # the final state (step 7 is not an output step)
    assert (tmp_path / "sync" / "output_000002.csv").read_bytes() != \
        (tmp_path / "sync" / "output_000004.csv").read_bytes()
    for f in H.FIELDS:
        assert np.ascontiguousarray(a.arr[f]).tobytes() == np.ascontiguousarray(b.arr[f]).tobytes(), f


@pytest.mark.parametrize("libname,name", [("lib3", "dam3d_small"), ("lib2", "dam2d_l0.01")])
def test_driver_matches_library_run(request, tmp_path, libname, name):
    """mps_amd/mps_sim3d (the reference main.c's flow: cal/params files,
    loader, initial parameters, simulation_run) writes the same snapshot
    bytes as the same entry points called through ctypes, and every
    snapshot is complete when the process exits."""
    import ctypes as C
    import subprocess
    import scenes
    lib3 = request.getfixturevalue(libname)     # the library of this dimension
    sc = scenes.CONFIGS[name]
    dim = sc["dim"]
    case = tmp_path / "case"
    case.mkdir()
    d = list(sc["domain"]) + [0.0] * (3 - dim)
    w = list(sc["water"]) + [0.0] * (3 - dim)
    dims = (C.c_double * 6)(*(d + w))
    n = lib3.mps_gen_dambreak(dims, sc["l0"], sc["wl"], sc["dl"], None, str(case / "particles.txt").encode())
    assert n > 0
    dt = 5e-4
    (case / "params.txt").write_text(
        "particle_distance %r\nwall_layers %d\ndummy_layers %d\ngravity_x 0\ngravity_y -9.81\ngravity_z 0\n"
        "dt %r\nt_end %r\noutput_interval 2\ncg_tolerance 1e-8\nmax_neighbors 300\n"
        "domain_x_min 0\ndomain_x_max %r\ndomain_y_min 0\ndomain_y_max %r\ndomain_z_min 0\ndomain_z_max %r\n"
        "output_dir out\n" % (sc["l0"], sc["wl"], sc["dl"], dt, 6.5 * dt, d[0], d[1], d[2]))
    (case / "cal.txt").write_text("particle_file particles.txt\nparam_file params.txt\n")
    import mps_amd
    driver = os.path.join(os.path.dirname(mps_amd.lib_path(dim)), "mps_sim%dd" % dim)
    say("driver", driver, "particles", n)
    p = subprocess.run([driver, "cal.txt"], cwd=str(case), capture_output=True, text=True, timeout=120)
    say(p.stdout[-3000:])
    say(p.stderr[-3000:])
    assert p.returncode == 0, p.stderr[-2000:]
    # the same flow through the library, into another directory
    abi = lib3.abi
    cfg = abi.SimConfig()
    lib3.config_set_defaults(C.byref(cfg))
    assert lib3.config_load_cal(str(case / "cal.txt").encode(), C.byref(cfg)) == 0
    assert lib3.config_load_params(cfg.param_file, C.byref(cfg)) == 0
    out_b = tmp_path / "lib_out"
    out_b.mkdir()
    cfg.output_dir = str(out_b).encode()
    lib3.set_config(cfg)
    lib3.set_option("tol_mode", 0)
    lib3.set_option("precision", 0)
    ps = lib3.particle_system_create(n)
    assert lib3.mps_load_particles(cfg.particle_file, ps) == 0
    lib3.particle_system_calc_initial_params(ps)
    say("simulation_run through ctypes")
    lib3.simulation_run(ps)
    assert lib3.mps_hip_last_error() == 0
    lib3.particle_system_free(ps)
    fa = sorted(os.listdir(case / "out"))
    assert fa == sorted(os.listdir(out_b)), (p.stdout[-1500:], p.stderr[-1500:])
    assert fa == sorted("output_%06d.%s" % (s, e) for s in (0, 2, 4, 6) for e in ("csv", "vtk"))
    assert sorted(os.listdir(out_b)) == fa
    for f in fa:
        x = (case / "out" / f).read_bytes()
        assert len(x) > 0 and x == (out_b / f).read_bytes(), f
