"""A GPU implicit solve recorded through the debug hook (the reference's ImplicitUpdateRecord,
TestGrid.cpp:398-453): header from the grid, one iterate per CR x update (ConjugateResiduals
calls its debug hook after each x update), and the last iterate equal to the grid's final
velocities, v = x + collision velocities."""
import numpy as np
import pytest

import mpmsnow_amd as G
from mpmsnow_amd.cr_record import ImplicitUpdateRecord, read_record

pytestmark = pytest.mark.gpu


def test_recorded_solve(tmp_path):
    print("[gpu-case] cr_record ball of 3000 particles on a moving plane", flush=True)
    rng = np.random.default_rng(5)
    n = 3000
    p = rng.uniform(-1, 1, (4 * n, 3))
    x = (np.float32([0.5, 0.2, 0.5]) + 0.1 * p[(p * p).sum(1) < 1][:n]).astype(np.float32)
    ctx = G.Context()
    ctx.particles_create(n)
    ctx.set("p", x)
    ctx.set("v", np.tile(np.float32([0.1, -1.0, 0.0]), (n, 1)))
    ctx.set("m", np.full(n, 1e-3, np.float32))
    ctx.model_snow(1.4e5, 0.2, 10.0, 2.5e-2, 7.5e-3)
    ctx.set_colliders([((0, 1, 0, -0.15), (0.2, 0.0, 0.0), 0.5, False)])
    h = 1.0 / 40
    g = G.Grid(ctx, np.arange(n), h)
    g.compute_particle_volumes()
    path = tmp_path / "debug.dat"
    with ImplicitUpdateRecord.for_grid(str(path), g, h, 0.002) as rec:
        st = g.update_grid_velocities(0.002, 200, 1e-5, debug=rec)
    r = read_record(str(path))
    print("[gpu-case] iterations %d, iterates %d" % (st["iterations"], len(r["iterates"])), flush=True)
    assert list(r["n"]) == [int(v) for v in g.n] and np.array_equal(r["min"], np.float32(g.min))
    assert r["h"] == np.float32(h)
    assert np.array_equal(r["masses"], g.get("masses").reshape(-1))
    assert len(r["iterates"]) == st["iterations"] > 1
    assert np.array_equal(r["iterates"][-1], g.get("velocities").reshape(-1))
    # the CR moves the explicit velocities towards the implicit ones: the first iterate differs
    assert not np.array_equal(r["iterates"][0], r["explicit_velocities"])
