"""Build recipe for the CPU oracle (test infrastructure, the checker).

Compiles oracle/mpm_oracle_capi.cpp into oracle/_build/libmpm_oracle.so with
g++ -O3 -fopenmp and -ffp-contract=off (no FMA contraction, no -march=native,
no -ffast-math), i.e. the arithmetic of a plain x86-64 build of the reference.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libmpm_oracle.so")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    src = os.path.join(HERE, "mpm_oracle_capi.cpp")
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
           "-Wall", "-Wno-unused", "-o", tmp, src]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
