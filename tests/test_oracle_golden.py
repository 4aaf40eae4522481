"""The CPU oracle reproduces the reference's golden vectors bit for bit.

The fixtures (tests/golden/*.npz + manifest.json) were written by
tests/golden/make_golden.py from the reference compiled from its own
sources; this test needs neither the reference nor a GPU.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import scenes
from mps_amd import abi as A
from oracle.oracle import Oracle, make_config

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(HERE, "manifest.json")) as f:
    MANIFEST = json.load(f)

FIELDS = ("pos", "vel", "acc", "pressure", "n", "type", "on_surface")


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def list_digest(start, idx):
    n = len(start) - 1
    counts = np.diff(start).astype(np.int32)
    return digest(np.concatenate([counts, idx.astype(np.int32)])) if n else ""


def test_fixture_files_match_manifest():
    for name, meta in MANIFEST.items():
        with open(os.path.join(HERE, name + ".npz"), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == meta["file_sha256"], name


@pytest.mark.parametrize("name", sorted(MANIFEST))
def test_oracle_reproduces_reference(name):
    meta = MANIFEST[name]
    gold = np.load(os.path.join(HERE, name + ".npz"))
    sc = scenes.CONFIGS[meta["scene"]]
    dim = sc["dim"]
    cfg = make_config(A.types(dim), **meta["params"])
    pos, vel, typ = scenes.dambreak(**sc)
    s = A.HostScene(A.types(dim), pos, vel, typ)
    orc = Oracle(dim)
    orc.init_params(s, cfg)
    assert s.ps.n0 == float(gold["n0"])
    assert s.ps.lambda_ == float(gold["lambda"])
    sim = orc.sim()
    sim.begin(s, cfg)
    st, ix = sim.neighbor_lists()
    assert list_digest(st, ix) == meta["digests"]["step0_lists"]
    for step in range(1, max(meta["checkpoints"]) + 1):
        sim.step(s, cfg)
        if step not in meta["checkpoints"]:
            continue
        for f in FIELDS:
            assert digest(s.arr[f]) == meta["digests"]["step%d_%s" % (step, f)], (name, step, f)
        st, ix = sim.neighbor_lists()
        assert list_digest(st, ix) == meta["digests"]["step%d_lists" % step], (name, step)
        assert np.array_equal(s.arr["pressure"], gold["step%d_pressure" % step])


def test_friction_fixture():
    """The oracle's hybrid wall friction (serial layout) reproduces the
    reference's own apply_friction on the stored state, bit for bit
    (golden/friction3ds.npz; 282 of its 320 fluid particles change)."""
    import harness as H
    sc, cfg, want, fl = H.friction_fixture(A.types(3, True))
    before = np.array(sc.arr["vel"])[fl]
    Oracle(3, True).apply_friction(sc, cfg)
    got = np.array(sc.arr["vel"])[fl]
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
    assert np.any(got != before)
