"""Builds the CPU oracle (test infrastructure) into oracle/libbsql_oracle.so.

g++ only: -O2 -ffp-contract=off -fno-fast-math so that every FP64 + - * /
rounds once, as IEEE-754 (and the HIP kernels) do.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "interp_oracle.cpp")
OUT = os.path.join(HERE, "libbsql_oracle.so")


def build(verbose=False):
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           "-fPIC", "-shared", "-pthread", "-Wall", "-o", OUT + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
