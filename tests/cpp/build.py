"""Build recipe of the C++ layer test programs (tests/cpp/_build/test_mpmsim, test_cook).

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


COOK_BIN = os.path.join(OUT_DIR, "test_cook")


def _compile(src, out):
    rpath = "$ORIGIN/" + os.path.relpath(LIB_DIR, OUT_DIR)
    tmp = out + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "mpmsnow_amd", "cpp"),
           os.path.join(HERE, src), "-L", LIB_DIR, "-lmpmsnow_hip",
           "-Wl,-rpath," + rpath, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for %s:\n%s" % (src, r.stderr))
    os.replace(tmp, out)


def build():
    """Both test programs: the MpmSim layer's (test_mpmsim) and the SOP cook driver's (test_cook)."""
    if not os.path.exists(os.path.join(LIB_DIR, "libmpmsnow_hip.so")):
        raise RuntimeError("build the GPU library first (mpmsnow_amd/build.py)")
    os.makedirs(OUT_DIR, exist_ok=True)
    _compile("test_mpmsim.cpp", BIN)
    _compile("test_cook.cpp", COOK_BIN)
    return BIN


if __name__ == "__main__":
    print(build())
