"""The MpmSim C++ class layer on the GPU: the test program (tests/cpp/test_mpmsim.cpp) runs the
reference's TestSimClass cases (initialization counts, three-step advance), TestGrid's splat
checks, the snow model's invariants, the host termination/debug hooks, a snowball landing
on a plane, all through MpmSim::Sim / Grid / SnowConstitutiveModel over the C-ABI, and
TestConjugateResiduals::testSolve through MpmSim::ConjugateResiduals over a ProceduralMatrix."""
import os
import subprocess

import pytest

from test_cpp_layer import load_recipe

pytestmark = pytest.mark.gpu


def test_cpp_layer_on_gpu():
    recipe = load_recipe()
    binary = recipe.BIN if os.path.exists(recipe.BIN) else recipe.build()
    print("[gpu-case] C++ layer test program", binary, flush=True)
    r = subprocess.run([binary], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    for name in ("testInitialization", "testTimestepAdvance", "testSplatting", "testSnowModel",
                 "testSolverHooks", "testSnowballOnGround", "testConjugateResidualsSolve"):
        assert "ok " + name in r.stdout
