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
