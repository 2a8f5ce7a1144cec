"""Build the fd23 HIP library (libfd23.so) in-tree for gfx950."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfd23.so")

SOURCES = ["fd23_ctx.hip", "fd23_box.hip", "fd23_material.hip", "host_precompute.cpp", "par.cpp", "fd23_serve.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed-FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose: bool = False) -> str:
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(HERE, "_build", os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + [
            "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"),
                   "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-c",
                   os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode())
            raise RuntimeError("compiling %s failed" % src)
        if verbose and out:
            sys.stderr.write(out.decode())
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    build_cli(verbose)
    return LIB


CLI_NAMES = ("nbpsv2d", "nbpsv2d_topabsorb", "nbpsv2d_exp", "nbpsv2d_exp_topabsorb", "sh2d", "sh2d_topabsorb",
             "psv2d_db", "psv2d_db_topabsorb")


def build_cli(verbose: bool = False) -> str:
    """The par-file driver, linked against libfd23.so, plus one link per
    reference program name (the name picks the variant)."""
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    exe = os.path.join(bindir, "fd23")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "fd23_cli.cpp"), "-o", exe + ".tmp", "-L", HERE, "-lfd23",
           "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(exe + ".tmp", exe)
    for name in CLI_NAMES:
        link = os.path.join(bindir, name)
        if os.path.lexists(link):
            os.remove(link)
        os.symlink("fd23", link)
    return exe


if __name__ == "__main__":
    print(build(verbose=True))
