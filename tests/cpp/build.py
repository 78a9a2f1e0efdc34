"""Build recipe of the MpmSim C++ layer test (tests/cpp/_build/test_mpmsim).

g++ compiles tests/cpp/test_mpmsim.cpp against the header-only layer
(mpmsnow_amd/cpp/MpmSim.hpp) and links the GPU library
(mpmsnow_amd/_build/libmpmsnow_hip.so), found at run time through an rpath
relative to the binary, so the tree can move.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT_DIR = os.path.join(HERE, "_build")
BIN = os.path.join(OUT_DIR, "test_mpmsim")
LIB_DIR = os.path.join(ROOT, "mpmsnow_amd", "_build")


def build():
    if not os.path.exists(os.path.join(LIB_DIR, "libmpmsnow_hip.so")):
        raise RuntimeError("build the GPU library first (mpmsnow_amd/build.py)")
    os.makedirs(OUT_DIR, exist_ok=True)
    rpath = "$ORIGIN/" + os.path.relpath(LIB_DIR, OUT_DIR)
    tmp = BIN + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "mpmsnow_amd", "cpp"),
           os.path.join(HERE, "test_mpmsim.cpp"), "-L", LIB_DIR, "-lmpmsnow_hip",
           "-Wl,-rpath," + rpath, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed:\n" + r.stderr)
    os.replace(tmp, BIN)
    return BIN


if __name__ == "__main__":
    print(build())
