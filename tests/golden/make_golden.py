#!/usr/bin/env python3
"""Writes golden vectors of the reference, computed by the reference itself.

Runs the reference compiled from its own sources (oracle/_ref, built by
`make -C oracle ref` where /root/reference exists) on small synthetic scenes
and stores, per checkpoint step: the full pressure and number-density
arrays, SHA-256 digests of every particle field and of the neighbour lists,
plus n0 and lambda.  These fixtures pin the CPU oracle (and through it the
GPU library) on machines where the reference is absent.

Serial OpenMP (OMP_NUM_THREADS=1) so the initial-parameter tie rule is the
reference's serial one (first strict maximum).

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

os.environ["OMP_NUM_THREADS"] = "1"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import scenes  # noqa: E402
from mps_amd import abi as A  # noqa: E402
from oracle.oracle import Reference, make_config  # noqa: E402

# name: (scene, variant, checkpoints, parameter overrides)
CASES = {
    "dam2d_cg": ("dam2d_l0.01", "2d", [1, 10, 100], dict(cg_tolerance=1e-6)),
    "shipped2d_iccg_hs": ("shipped2d", "2d", [1, 10], dict(solver_type=1, hs_mode=1, influence_ratio_lap=4.0,
                                                          cg_tolerance=1e-8)),
    "dam3d_small_cg": ("dam3d_small", "3dp", [1, 5], dict(cg_tolerance=1e-6)),
    "shipped3d_iccg_hs": ("shipped3d", "3dp", [1], dict(solver_type=1, hs_mode=1, influence_ratio_lap=2.1,
                                                        cg_tolerance=1e-8)),
    "shipped3d_serial_cg": ("shipped3d", "3d", [1, 3], dict(cg_tolerance=1e-6, cmps_gradient=1)),
}

FIELDS = ("pos", "vel", "acc", "pressure", "n", "type", "on_surface")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_case(name):
    scene, variant, checkpoints, over = CASES[name]
    sc = scenes.CONFIGS[scene]
    dim = sc["dim"]
    params = scenes.params_for(scene, **over)
    ref = Reference(variant)
    cfg = ref.config_from(make_config(A.types(dim), **params))
    ref.set_config(cfg)
    pos, vel, typ = scenes.dambreak(**sc)
    s = A.HostScene(A.types(dim), pos, vel, typ)
    ref.L.particle_system_calc_initial_params(s.ptr())
    n = s.num
    nl, cl = ref.make_lists(n)
    ref.L.neighbor_search_cell_linked_list(nl, s.ptr(), cl, ref.search_radius())
    out = {"n0": np.float64(s.ps.n0), "lambda": np.float64(s.ps.lambda_)}
    meta = {"scene": scene, "variant": variant, "params": params, "checkpoints": checkpoints,
            "n": n, "digests": {}}
    lists = A.padded_lists(nl.contents, n)
    meta["digests"]["step0_lists"] = digest(np.concatenate([np.asarray([len(x)], np.int32) for x in lists] +
                                                           [np.concatenate(lists)]))
    for step in range(1, max(checkpoints) + 1):
        ref.L.simulation_step(s.ptr(), nl, cl, step)
        if step in checkpoints:
            a = s.arr
            for f in FIELDS:
                meta["digests"]["step%d_%s" % (step, f)] = digest(a[f])
            lists = A.padded_lists(nl.contents, n)
            meta["digests"]["step%d_lists" % step] = digest(
                np.concatenate([np.asarray([len(x)], np.int32) for x in lists] + [np.concatenate(lists)]))
            out["step%d_pressure" % step] = np.ascontiguousarray(a["pressure"])
            out["step%d_n" % step] = np.ascontiguousarray(a["n"])
            out["step%d_pos" % step] = np.ascontiguousarray(a["pos"])
    ref.free_lists(nl, cl)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    return meta, path


def main():
    if not Reference.available("3dp"):
        sys.exit("oracle/_ref not built: run make -C oracle ref where the reference sources exist")
    manifest = {}
    for name in CASES:
        meta, path = run_case(name)
        with open(path, "rb") as f:
            meta["file_sha256"] = hashlib.sha256(f.read()).hexdigest()
        manifest[name] = meta
        print(name, "n=%d" % meta["n"], os.path.getsize(path), "bytes")
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
