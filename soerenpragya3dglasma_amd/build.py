"""Build recipe of libglasma.so (HIP, gfx950) and the host-only algebra check library.

``python -m soerenpragya3dglasma_amd.build`` compiles every ``csrc/*.hip`` with hipcc for gfx950 and links
them into ``soerenpragya3dglasma_amd/_lib/libglasma.so`` (in-tree, so the built library travels with the
source tree).  The host C++ drivers (``csrc/*.cpp``) are linked into the same library, and
``csrc/glasma_solver_main.cpp`` becomes the reference-style command line ``_lib/glasma_solver``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
OBJ = os.path.join(OUT, "obj")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GLASMA_ARCH", "gfx950")
# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wno-unused-result", "-Wno-unused-value",
]
HOST_CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-I", os.path.join(ROOT, "include"), "-I", CSRC]

LIB_NAME = "libglasma.so"
# extra compiler flags for experiments (e.g. -DGLASMA_NT_STORES) and an alternative output directory
EXTRA = os.environ.get("GLASMA_CXXFLAGS", "").split()


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp") and f not in ("glasma_solver_main.cpp", "hostcheck.cpp"))
    return hip, cpp


def build(verbose: bool = False, jobs: int = 8, out_dir: str | None = None) -> str:
    out = os.path.abspath(out_dir) if out_dir else OUT
    obj = os.path.join(out, "obj")
    os.makedirs(obj, exist_ok=True)
    flags = HIP_FLAGS + EXTRA
    hip, cpp = _sources()
    cmds = []
    objs = []
    for f in hip:
        o = os.path.join(obj, f + ".o")
        objs.append(o)
        cmds.append([HIPCC, *flags, "-c", os.path.join(CSRC, f), "-o", o])
    for f in cpp:
        o = os.path.join(obj, f + ".o")
        objs.append(o)
        cmds.append([HIPCC, *flags, "-x", "hip", "-c", os.path.join(CSRC, f), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for c, o in zip(cmds, outs):
            print(" ".join(c))
            if o.strip():
                print(o)
    lib_tmp = os.path.join(out, LIB_NAME + ".tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", lib_tmp, *objs])
    lib = os.path.join(out, LIB_NAME)
    os.replace(lib_tmp, lib)
    main = os.path.join(CSRC, "glasma_solver_main.cpp")
    if os.path.exists(main):
        _run([HIPCC, *flags, "-x", "hip", main, "-o", os.path.join(out, "glasma_solver"),
              "-L", out, "-lglasma", "-Wl,-rpath,$ORIGIN"])
    build_hostcheck(out)
    return lib


def build_hostcheck(out: str | None = None) -> str:
    """Host (g++) build of the per-site device code, for CPU tests against the oracle."""
    out = os.path.abspath(out) if out else OUT
    os.makedirs(out, exist_ok=True)
    path = os.path.join(out, "libglasma_hostcheck.so")
    if shutil.which("g++"):
        _run(["g++", *HOST_CXX_FLAGS, "-shared", "-o", path, os.path.join(CSRC, "hostcheck.cpp")])
    return path


if __name__ == "__main__":
    od = None
    if "--out" in sys.argv:
        od = sys.argv[sys.argv.index("--out") + 1]
    path = build(verbose="-v" in sys.argv, out_dir=od)
    print(path)
