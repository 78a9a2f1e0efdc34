"""The oracle against the committed golden fixtures (CPU): regenerating each fixture from its seeded
inputs (tests/golden/make_golden.py) reproduces the stored FP64 outputs to 1e-12 of each array's
magnitude, and the integer outputs (grid extent, collision flags, iteration counts) exactly. A
change to the oracle that moves any of them fails here before it reaches the GPU parity tests."""
import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
make_golden = importlib.util.module_from_spec(spec)
spec.loader.exec_module(make_golden)


def load(name):
    return np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("name", sorted(make_golden.MAKERS))
def test_oracle_reproduces_golden_fixture(oracle, name):
    ref = load(name)
    new = make_golden.MAKERS[name](oracle)
    assert sorted(ref.files) == sorted(k for k, v in new.items() if not isinstance(v, dict))
    for k in ref.files:
        a, b = np.asarray(new[k]), ref[k]
        assert a.shape == b.shape, k
        if np.issubdtype(b.dtype, np.integer):
            assert np.array_equal(a, b), k
        else:
            scale = max(np.abs(b).max(), 1e-300)
            assert np.abs(a.astype(np.float64) - b).max() <= 1e-12 * scale, k


def test_fixtures_are_small():
    total = sum(os.path.getsize(os.path.join(HERE, "golden", n + ".npz")) for n in make_golden.MAKERS)
    assert total < 400 * 1024
