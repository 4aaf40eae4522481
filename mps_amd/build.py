"""Build recipe of libmps_hip{3,2}d.so, libmps_hip3ds.so and the mps_sim
drivers.  libmps_hip3ds.so is the 3-D library in the layout of the 3-D
serial program (mps_fluid3D: SimConfig with the hybrid-friction fields,
MPS_FRICTION_FIELDS), with its wall friction.

HIP sources are compiled by hipcc for gfx950 only, FP64 contraction off (the
reference's x86 build has no FMA and parity is bit-exact), packed-FP32 ops
disabled; host C sources by gcc with OpenMP.  Everything is written in-tree
under mps_amd/ so the built libraries travel with the repository snapshot.

    python -m mps_amd.build            # 3d, 2d and 3ds
    python -m mps_amd.build --dim 3
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
HIP_SRC = ["k_particles.hip", "k_pack.hip", "k_cg.hip", "k_ic.hip", "k_extra.hip", "k_adjoint.hip", "runtime.hip", "dropins.hip"]
C_SRC = ["config.c", "particles.c", "io.c"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-fvisibility=hidden",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-parameter"]
C_FLAGS = ["-O2", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fvisibility=hidden", "-Wall", "-Wextra",
           "-Wno-unused-parameter", "-Wno-stringop-truncation"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s%s" % (" ".join(cmd), p.stdout, p.stderr))
    return p.stdout + p.stderr


def lib_path(dim, serial=False):
    return os.path.join(HERE, "libmps_hip%dd%s.so" % (dim, "s" if serial else ""))


def build_dim(dim, jobs=8, verbose=False, serial=False):
    tag = "%dd%s" % (dim, "s" if serial else "")
    bdir = os.path.join(HERE, "_build", tag)
    os.makedirs(bdir, exist_ok=True)
    defs = ["-DMPS_DIM=%d" % dim, "-I" + INC, "-I" + os.path.join(HERE, "csrc")]
    if serial:
        defs.append("-DMPS_FRICTION_FIELDS")
    cmds = []
    objs = []
    for s in HIP_SRC:
        src = os.path.join(HERE, "csrc", s)
        if not os.path.exists(src):
            continue
        o = os.path.join(bdir, s + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + defs + ["-c", src, "-o", o])
        objs.append(o)
    for s in C_SRC:
        src = os.path.join(HERE, "host", s)
        o = os.path.join(bdir, s + ".o")
        cmds.append(["gcc"] + C_FLAGS + defs + ["-c", src, "-o", o])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    gomp = subprocess.run(["gcc", "-print-file-name=libgomp.so"], capture_output=True,
                          text=True).stdout.strip()
    tmp = lib_path(dim, serial) + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs +
         [gomp, "-pthread", "-Wl,--dynamic-list=" + os.path.join(HERE, "csrc", "dynamic.list")])
    os.replace(tmp, lib_path(dim, serial))
    # stand-alone driver, same CLI as the reference's mps_sim
    exe = os.path.join(HERE, "mps_sim%s" % tag)
    _run(["gcc", "-O2", "-Wall"] + [x for x in defs if x.startswith("-D")] + ["-I" + INC,
          os.path.join(HERE, "host", "main.c"), "-o", exe, "-L" + HERE,
          "-l:libmps_hip%s.so" % tag, "-Wl,-rpath,$ORIGIN"])
    return lib_path(dim, serial)


def build(dims=(3, 2), jobs=8, verbose=False, serial3=True):
    out = []
    for d in dims:
        out.append(build_dim(d, jobs=jobs, verbose=verbose))
    if serial3 and 3 in dims:
        out.append(build_dim(3, jobs=jobs, verbose=verbose, serial=True))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dim", type=int, choices=(2, 3), action="append")
    ap.add_argument("-j", "--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    for p in build(tuple(a.dim or (3, 2)), jobs=a.jobs, verbose=a.verbose):
        print("built", p)


if __name__ == "__main__":
    main()
