#!/usr/bin/env python3
"""Stores the reference's own shipped trajectory of its 3-D dam break as a
fixture: the snapshots mps_fluid3D_parallel/output/output_000100.csv and
output_000200.csv, written by the reference's output_csv (io.c:10-34,
"%.8e" fields) from examples/dambreak (28,112 particles; params.txt: ICCG,
hs_mode 1, clamp 1, cg_tolerance 1e-8 absolute, influence ratios 2.1).

Per snapshot: the positions of the fluid particles (the walls do not move),
every particle's pressure and type, as the float64 values the text reads
to.  Data only: nothing of the reference's code is kept.

    python tests/golden/make_shipped_output_fixture.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/mps_fluid3D_parallel/output"
STEPS = (100, 200)


def main():
    if not os.path.isdir(SRC):
        sys.exit("reference outputs absent: " + SRC)
    out = {}
    init = np.loadtxt(os.path.join(SRC, "output_000000.csv"), delimiter=",", skiprows=1)
    fluid = np.nonzero(init[:, 7].astype(int) == 0)[0].astype(np.int32)
    out["fluid_index"] = fluid
    for s in STEPS:
        a = np.loadtxt(os.path.join(SRC, "output_%06d.csv" % s), delimiter=",", skiprows=1)
        assert a.shape[0] == init.shape[0], "snapshot %d drops particles" % s
        out["step%d_fluid_pos" % s] = a[fluid, 0:3].copy()
        out["step%d_pressure" % s] = a[:, 6].copy()
        out["step%d_type" % s] = a[:, 7].astype(np.int8)
    path = os.path.join(HERE, "shipped3d_output.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes,", len(fluid), "fluid particles")


if __name__ == "__main__":
    main()
