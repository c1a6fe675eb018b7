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
CXX_SOURCES = ["program.cpp", "parser.cpp", "capi.cpp", "learned.cpp", "table.cpp", "steps.cpp", "regex.cpp",
               "kernel_gen.cpp"]
# device headers built into the library for kernels generated at run time,
# by the names the sources include them with
GEN_HEADERS = [("vm_device.h", os.path.join(CSRC, "vm_device.h")), ("devplan.h", os.path.join(CSRC, "devplan.h")),
               ("pow10.h", os.path.join(CSRC, "pow10.h")),
               ("../../include/bsql/interp.h", os.path.join(HERE, "..", "include", "bsql", "interp.h"))]

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


def _write_gen_headers():
    """gen_headers.inc: the device headers' texts (system includes removed;
    hipRTC supplies the HIP runtime) for kernel_gen.cpp."""
    names, texts = [], []
    for name, path in GEN_HEADERS:
        lines = [l for l in open(path).read().splitlines() if not l.startswith("#include <")]
        text = "\n".join(lines) + "\n"
        assert ")BSQLGEN\"" not in text
        names.append(name)
        texts.append(text)
    out = ["static const int kGenHeaderCount = %d;" % len(names),
           "static const char* kGenHeaderNames[] = {%s};" % ", ".join('"%s"' % n for n in names),
           "static const char* kGenHeaderTexts[] = {"]
    for t in texts:
        # one raw literal per piece keeps each under compiler string limits
        pieces = [t[i:i + 60000] for i in range(0, len(t), 60000)]
        out.append("    " + " ".join('R"BSQLGEN(%s)BSQLGEN"' % p for p in pieces) + ",")
    out.append("};")
    path = os.path.join(BUILD, "gen_headers.inc")
    body = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != body:
        with open(path, "w") as f:
            f.write(body)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    _write_gen_headers()
    cmds, objs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmds.append(["g++"] + CXX_FLAGS + ["-I" + BUILD, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run_all(cmds, verbose)
    tmp = OUT + ".tmp"
    _run_all([[HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
              ["-L/opt/rocm/lib", "-lhiprtc"]], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
