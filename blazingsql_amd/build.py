"""Builds libbsql.so (C ABI + gfx950 kernels) in-tree.

hipcc cross-compiles the device code for gfx950 without a GPU; the host
C++ parts are compiled with g++. Every FP64 + - * / must round once, so all
code is built with -ffp-contract=off.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libbsql.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_INC = "/opt/rocm/include"

HIP_SOURCES = ["interp_kernels.hip", "learned.hip", "strfn.hip", "noalloc.hip"]
CXX_SOURCES = ["program.cpp", "parser.cpp", "capi.cpp", "learned.cpp", "table.cpp", "steps.cpp", "regex.cpp"]

HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
             "-D__HIP_PLATFORM_AMD__", "-I" + ROCM_INC]


def _run_all(cmds, verbose):
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the device-only feature flag
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l]
        if p.returncode != 0 or (verbose and lines):
            print("\n".join(lines), file=sys.stderr)
        if p.returncode != 0:
            failed.append(cmd)
    if failed:
        raise RuntimeError("build failed: " + " ".join(failed[0]))


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    cmds, objs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run_all(cmds, verbose)
    tmp = OUT + ".tmp"
    _run_all([[HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
