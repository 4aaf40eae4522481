"""The particles.txt loader (mps_load_particles, which parses batches of
lines in parallel) against the oracle's line-by-line restatement of the
reference's load_particles (mps_fluid3D_parallel/src/main.c:20-85; 2-D
mps_fluid2D/src/main.c:20-85): the same particles, bit for bit, and the same
messages on stdout and stderr, over files built to exercise every rule of
the fgets/sscanf loop."""
import ctypes as C
import os

import numpy as np
import pytest

from mps_amd import abi as A
from oracle.oracle import Oracle

_libc = C.CDLL(None)


def _flush():
    _libc.fflush(None)


def _empty_system(abi, cap):
    arr = (abi.Particle * max(cap, 1))()
    ps = abi.ParticleSystem()
    ps.particles = C.cast(arr, C.POINTER(abi.Particle))
    ps.num = 0
    ps.capacity = cap
    return ps, arr


def _rows(dim, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, (n, 2 * dim))
    t = rng.integers(0, 3, n)
    fmt = " ".join(["%.8e"] * (2 * dim)) + " %d\n"
    return [fmt % (tuple(x[i]) + (t[i],)) for i in range(n)]


def _adversarial(dim, seed):
    good = _rows(dim, 40, seed)
    cols = 2 * dim
    lines = ["# leading comment\n", "\n", "   \t# indented comment\n", "\r\n",
             "  37 particles follow\n"]                       # the count line: sscanf %d reads 37
    lines += good[:3]
    lines += ["\t" + good[3].replace(" ", "\t"),               # tabs between fields
              good[4].replace("\n", "\r\n"),                   # CRLF
              "# a comment between rows\n",
              "\n",
              " ".join(["1.0"] * (cols - 1)) + " abc\n",       # malformed: last field not an int
              "1.0 2.0\n",                                     # malformed: too few fields
              "\f\n",                                          # sscanf returns EOF
              good[5].rstrip("\n") + " trailing tokens 9 9\n",  # extra tokens are ignored
              " ".join(["0.5"] * cols) + " 1.7\n",             # %d reads the 1 of 1.7
              " " * 600 + good[6],                             # > 511 chars: split; tail parsed
              "#" + "x" * 520 + " 1 2 3\n",                    # long comment: its tail is malformed
              " " * 505 + good[7]]                             # split inside the row
    lines += good[8:40]                                        # more rows than announced
    text = "".join(lines)
    return text.rstrip("\n")                                   # no final newline


def _load_both(dim, path, cap, capfd):
    import mps_amd
    lib = mps_amd.load(dim)
    orc = Oracle(dim)
    abi = A.types(dim)
    _flush()              # C stdio output of earlier tests is not ours
    capfd.readouterr()
    ps_l = lib.particle_system_create(cap)
    rc_l = lib.mps_load_particles(str(path).encode(), ps_l)
    _flush()
    out_l = capfd.readouterr()
    ps_o, keep = _empty_system(abi, cap)
    rc_o = orc.L.orc_load_particles(str(path).encode(), C.byref(ps_o))
    _flush()
    out_o = capfd.readouterr()
    n_l, n_o = ps_l.contents.num, ps_o.num
    a = C.string_at(ps_l.contents.particles, C.sizeof(abi.Particle) * n_l)
    b = C.string_at(ps_o.particles, C.sizeof(abi.Particle) * n_o)
    lib.particle_system_free(ps_l)
    del keep
    return (rc_l, n_l, a, out_l.out, out_l.err), (rc_o, n_o, b, out_o.out, out_o.err)


@pytest.mark.parametrize("dim", [3, 2])
@pytest.mark.parametrize("cap", [64, 20])
def test_loader_rules_match_reference_loop(dim, cap, tmp_path, capfd):
    path = tmp_path / "particles.txt"
    path.write_text(_adversarial(dim, seed=dim))
    got, want = _load_both(dim, path, cap, capfd)
    assert got == want
    rc, n, _, out, err = got
    assert rc == 0 and out.startswith("Loaded particles:")
    assert err.count("skipping malformed line") >= 4
    if cap == 20:
        assert n == 20 and "capacity exceeded" in err


@pytest.mark.parametrize("dim", [3, 2])
def test_loader_fewer_rows_than_announced(dim, tmp_path, capfd):
    path = tmp_path / "particles.txt"
    path.write_text("50\n" + "".join(_rows(dim, 30, 5)))
    got, want = _load_both(dim, path, 64, capfd)
    assert got == want
    assert got[1] == 30 and "expected 50 particles but loaded 30" in got[4]


def test_loader_batches(tmp_path, capfd):
    """Rows across several parallel batches, with malformed and comment
    lines on both sides of a batch boundary."""
    rows = _rows(3, 140000, 11)
    for at in (65534, 65535, 65536, 65537, 131071, 131073):
        rows[at] = ("# c\n" if at % 2 else "1.0 x\n")
    path = tmp_path / "particles.txt"
    path.write_text("%d\n" % len(rows) + "".join(rows))
    got, want = _load_both(3, path, len(rows), capfd)
    assert got == want
    assert got[1] == len(rows) - 6


def test_loader_errors(tmp_path, capfd):
    missing = tmp_path / "none.txt"
    got, want = _load_both(3, missing, 8, capfd)
    assert got == want and got[0] == -1
    bad = tmp_path / "bad.txt"
    bad.write_text("# only comments\n\nnot a count\n")
    got, want = _load_both(3, bad, 8, capfd)
    assert got == want and got[0] == -1 and "invalid particle count" in got[4]


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="reference inputs absent")
@pytest.mark.parametrize("rel,dim", [("mps_fluid3D_parallel/examples/dambreak/particles.txt", 3),
                                     ("mps_fluid2D/examples/dambreak/particles.txt", 2)])
def test_oracle_loader_reads_shipped_inputs(rel, dim, capfd):
    """The oracle's loader on the reference's own input files gives the
    values Python's correctly rounded float() reads from the same text."""
    import scenes
    path = os.path.join("/root/reference", rel)
    pos, vel, typ = scenes.load_particles_txt(path, dim)
    abi = A.types(dim)
    ps, keep = _empty_system(abi, len(typ))
    _flush()
    capfd.readouterr()
    assert Oracle(dim).L.orc_load_particles(path.encode(), C.byref(ps)) == 0
    _flush()
    out = capfd.readouterr().out
    arr = np.ctypeslib.as_array(ps.particles, shape=(ps.num,))
    assert ps.num == len(typ)
    assert np.array_equal(arr["pos"], pos) and np.array_equal(arr["vel"], vel)
    assert np.array_equal(arr["type"], typ)
    assert out == "Loaded particles: %d fluid, %d wall, %d total\n" % (
        (typ == 0).sum(), (typ == 1).sum(), len(typ))
