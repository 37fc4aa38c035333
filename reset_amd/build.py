"""Builds libreset_amd.so (all HIP sources, gfx950) and the CLI tools in-tree.

    python -m reset_amd.build            # library + tools
    python -m reset_amd.build --lib-only

Every call recompiles (object files are cached by content hash of the source
and its headers only within one call's build folder).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libreset_amd.so")
TOOLS_DIR = os.path.join(HERE, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 ops (v_pk_*_f32) are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}",
]

SOURCES = ["core.hip", "tables.hip", "tc.hip", "pairs.hip", "sss.hip", "sss_join.hip", "sss_group.hip", "ir.hip",
           "gen.hip", "autotune.hip", "solver.hip", "competitors.hip"]
TOOLS = ["tc", "sss", "ir", "intersection"]
# the standalone apps' make_hash_table, one object per app (their globals differ)
STANDALONE_DIR = os.path.join(HERE, "standalone")
STANDALONE_APPS = {"tc": "RESET_AMD_APP_TC", "sss": "RESET_AMD_APP_SSS", "ir": "RESET_AMD_APP_IR"}


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_lib(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_tools(jobs: int = 4) -> list:
    os.makedirs(TOOLS_DIR, exist_ok=True)
    cli = os.path.join(CSRC, "cli")
    todo = [t for t in TOOLS if os.path.exists(os.path.join(cli, t + ".cpp"))]

    def one(t):
        out = os.path.join(TOOLS_DIR, t)
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", os.path.join(cli, t + ".cpp"), "-o", out,
              f"-L{HERE}", "-lreset_amd", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/.."])
        return out

    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(todo) or 1))) as ex:
        out = list(ex.map(one, todo))
    # make_hash_table objects, and a TC main in the free-function form linked with one
    os.makedirs(STANDALONE_DIR, exist_ok=True)
    src = os.path.join(CSRC, "standalone", "make_hash_table.cpp")

    def obj(app):
        o = os.path.join(STANDALONE_DIR, f"{app}_hashtable.o")
        _run([HIPCC, *HIP_FLAGS, "-x", "hip", f"-D{STANDALONE_APPS[app]}", "-c", src, "-o", o])
        return o

    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        objs = dict(zip(STANDALONE_APPS, ex.map(obj, STANDALONE_APPS)))
    exe = os.path.join(TOOLS_DIR, "standalone_tc")
    _run([HIPCC, *HIP_FLAGS, "-x", "hip", os.path.join(cli, "standalone_tc.cpp"), "-x", "none", objs["tc"],
          "-o", exe, f"-L{HERE}", "-lreset_amd", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/.."])
    return out + [exe]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    print(build_lib(a.j))
    if not a.lib_only:
        for t in build_tools():
            print(t)


if __name__ == "__main__":
    sys.exit(main())
