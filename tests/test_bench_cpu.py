"""bench.py without a GPU: the byte models that feed the roofline figures,
checked against values worked out by hand, and the one-GPU contract."""
import os
import subprocess
import sys
from types import SimpleNamespace

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _st(z, ne, pairs=0):
    return SimpleNamespace(nnz=z, n_eq=ne, n_pairs=pairs)


def test_spmv_bytes_csr_formula():
    # Z (8 + 4) + 4 (n_eq + 1) + 2 * 8 * n_eq for Z = 10, n_eq = 3
    assert bench.spmv_bytes(_st(10, 3)) == 120 + 16 + 48
    # FP32 values and vectors: Z (4 + 4) + 4 (n_eq + 1) + 2 * 4 * n_eq
    assert bench.spmv_bytes(_st(10, 3), 4, 4) == 80 + 16 + 24


def test_coded_bytes():
    # one 4-byte word per off-diagonal (Z - n_eq of them), 20 B per row
    assert bench.coded_bytes(_st(10, 3)) == 4 * 7 + 3 * 20


def test_step_bytes_config1():
    # the 1,077,505-particle dam break at the step the bench times:
    # 118,619,656 pairs, 88,034,345 nonzeros, 835,671 unknowns, 523.7
    # iterations -> 4.62 GB of particle phases + 1.12 GB per iteration
    st = _st(88034345, 835671, 118619656)
    b_part = 28 * 118619656 + 88034345 * 12 + 4 * 835672 + 8 * 835671 + 216 * 1077505
    b_cg = 88034345 * 12 + 4 * 835672 + 72 * 835671
    assert bench.step_bytes(st, 1077505, 523.7) == b_part + 523.7 * b_cg
    assert abs(bench.step_bytes(st, 1077505, 523.7) / 8e12 - 0.0739) < 5e-4   # ~73.9 ms at 8 TB/s


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in p.stderr
