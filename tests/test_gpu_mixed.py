"""Mixed precision PPE solves (option "precision" 1: FP32 matrix with FP64
CG vectors; 2: FP32 matrix and vectors; 3: FP32 vectors, off-diagonals
from 16-bit distance codes) against the FP64 GPU solve of the same scene.

Nothing in the reference pins the mixed form; its contract is the FP64
solution: relative L2 pressure error <= 1e-6 at every step, positions within
the same bound, and the solve must stop on the FP64 residual (so its
iteration count stays close to the FP64 count).
"""
import sys

import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", [1, 2, 3])
@pytest.mark.parametrize("name,steps", [("dam3d_small", 6), ("dam3d_mid", 3)])
def test_mixed_matches_fp64(lib3, name, steps, precision):
    dim, cfg, pos, vel, typ = H.build(name, cg_tolerance=1e-9)
    lib3.set_option("tol_mode", 1)
    say("mixed (precision %d) vs fp64" % precision, name, "N=%d" % len(typ), "steps", steps)
    lib3.set_option("precision", 0)
    g64 = H.GpuRun(lib3, cfg, pos, vel, typ)
    n0, lam = g64.scene.ps.n0, g64.scene.ps.lambda_
    gmx = H.GpuRun(lib3, cfg, pos, vel, typ, n0=n0, lam=lam)
    try:
        for s in range(1, steps + 1):
            lib3.set_option("precision", 0)
            assert g64.step() == 0
            st64 = lib3.stats(g64.scene)
            lib3.set_option("precision", precision)
            assert gmx.step() == 0
            stmx = lib3.stats(gmx.scene)
            a, b = gmx.scene.arr, g64.scene.arr
            ep = H.rel_err(a["pressure"], b["pressure"])
            ex = H.rel_err(a["pos"], b["pos"])
            say("step %d: p err %.3e, pos err %.3e, iters fp64 %d mixed %d (reliable updates %d), "
                "res %.3e / bnorm %.3e" % (s, ep, ex, st64.last_iters, stmx.last_iters,
                                           stmx.last_outer, stmx.last_residual, stmx.last_bnorm))
            assert ep <= 1e-6
            assert ex <= 1e-6
            # converged on the FP64 residual
            assert stmx.last_residual <= 1e-9 * stmx.last_bnorm * (1 + 1e-12)
            assert stmx.last_iters <= 1.5 * st64.last_iters + 20
            assert np.array_equal(a["on_surface"], b["on_surface"]) or s > 1
    finally:
        lib3.set_option("precision", 0)
        g64.close()
        gmx.close()
