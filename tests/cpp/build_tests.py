"""Builds the C++ operator-surface test (tests/cpp/test_kernels) against the
in-tree libbsql.so; the GPU tests run the binary."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BINARY = os.path.join(HERE, "test_kernels")


def build(verbose=False):
    cmd = ["/opt/rocm/bin/hipcc", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "test_kernels.cpp"), "-L", os.path.join(ROOT, "blazingsql_amd"), "-lbsql",
           "-Wl,-rpath,$ORIGIN/../../blazingsql_amd", "-o", BINARY]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return BINARY


if __name__ == "__main__":
    build(verbose=True)
