"""mps_hip_solve_csr_noalloc, the CSR solve whose device buffers all live in
one static array, against mps_hip_solve_csr through the C-ABI: the same
solution bit for bit and the same iteration count, on a 3-D PPE (the
dam3d_small scene's first pressure system) and then on a smaller 2-D
Laplacian, CG and ICCG; and a solve above the array's size is refused
before anything is written."""
import ctypes as C

import numpy as np
import pytest

import harness as H

pytestmark = pytest.mark.gpu


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def _laplacian(nx, ny, shift=0.05):
    rows, cols, vals = [], [], []
    for i in range(nx):
        for j in range(ny):
            k = i * ny + j
            for di, dj in ((-1, 0), (0, -1), (0, 0), (0, 1), (1, 0)):
                a, b = i + di, j + dj
                if 0 <= a < nx and 0 <= b < ny:
                    rows.append(k)
                    cols.append(a * ny + b)
                    vals.append(4.0 + shift if (di, dj) == (0, 0) else -1.0)
    n = nx * ny
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, np.array(rows) + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    order = np.lexsort((cols, rows))
    return rp, np.array(cols, np.int32)[order], np.array(vals)[order]


def _solve(lib, entry, rp, col, val, b, solver, tol):
    n = len(rp) - 1
    x = np.full(n, np.nan)
    it = C.c_int(-1)
    fn = getattr(lib, entry)
    rc = fn(n, _p(rp, C.c_int), _p(col, C.c_int), _p(val), _p(b), _p(x), solver, 5000, tol, C.byref(it))
    return rc, x, it.value


@pytest.fixture(scope="module")
def ppe3d(lib3):
    dim, cfg, pos, vel, typ = H.build("dam3d_small", cg_tolerance=1e-10)
    lib3.set_option("tol_mode", 1)
    g = H.GpuRun(lib3, cfg, pos, vel, typ)
    assert g.step() == 0
    q = lib3.export_ppe(g.scene)
    g.close()
    return q["row_ptr"], q["col"], q["val"], q["b"]


@pytest.mark.parametrize("solver", [0, 1])
def test_noalloc_equals_solve_csr(lib3, ppe3d, solver):
    lib3.set_option("tol_mode", 1)
    big = ppe3d
    rng = np.random.default_rng(11)
    rp2, col2, val2 = _laplacian(31, 23)
    small = (rp2, col2, val2, rng.standard_normal(len(rp2) - 1))
    for name, (rp, col, val, b) in (("3-D PPE", big), ("2-D Laplacian", small), ("3-D PPE again", big)):
        H.say("solver", solver, name, "unknowns", len(rp) - 1, "entries", int(rp[-1]))
        rc0, x0, it0 = _solve(lib3, "mps_hip_solve_csr", rp, col, val, b, solver, 1e-10)
        rc1, x1, it1 = _solve(lib3, "mps_hip_solve_csr_noalloc", rp, col, val, b, solver, 1e-10)
        assert rc0 == 0 and rc1 == 0, (rc0, rc1, lib3.mps_hip_error_string(rc1))
        assert it0 == it1 and 0 < it1 < 5000
        assert np.array_equal(x0.view(np.uint64), x1.view(np.uint64))
        H.say("  iterations", it1, "bit-identical")


def test_noalloc_rejects_above_scratch(lib3):
    """800,000 unknowns need more than the 96 MiB array even as a diagonal
    matrix (about 200 bytes per unknown): refused, x untouched."""
    n = 800000
    rp = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    val = np.full(n, 2.0)
    b = np.ones(n)
    rc, x, it = _solve(lib3, "mps_hip_solve_csr_noalloc", rp, col, val, b, 0, 1e-10)
    assert rc == 5  # MPS_ERR_ARG
    msg = lib3.mps_hip_error_string(rc).decode()
    assert "scratch" in msg, msg
    assert np.all(np.isnan(x)) and it == 0
    # and the same call through the allocating entry point still works
    rc, x, it = _solve(lib3, "mps_hip_solve_csr", rp, col, val, b, 0, 1e-10)
    assert rc == 0 and np.allclose(x, 0.5)
