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


def test_recorded_iterates_against_the_oracle(oracle):
    """The iterate stream itself against the oracle's: the same solve in the FP32 and FP64 oracle
    with the reference's debug hook (LinearSolver::Debug, TestGrid.cpp:398-453). The GPU records
    as many iterates as the FP32 oracle (one per x update, +-1 where |r|^2 ends within rounding of
    the threshold), and every iterate k lies within the FP32 oracle's own distance from the FP64
    iterate k (3x, floor 1e-5 of the solution's size)."""
    print("[gpu-case] cr iterates vs oracle: ball of 2000 particles on a moving plane", flush=True)
    rng = np.random.default_rng(6)
    n = 2000
    p = rng.uniform(-1, 1, (4 * n, 3))
    x = (np.float32([0.5, 0.2, 0.5]) + 0.1 * p[(p * p).sum(1) < 1][:n]).astype(np.float32)
    v = np.tile(np.float32([0.1, -1.0, 0.0]), (n, 1)) + rng.uniform(-0.2, 0.2, (n, 3)).astype(np.float32)
    m = np.full(n, 1e-3, np.float32)
    h, dt = 1.0 / 40, 0.002
    plane = ((0, 1, 0, -0.15), (0.2, 0.0, 0.0), 0.5, False)
    ctx = G.Context()
    ctx.particles_create(n)
    ctx.set("p", x)
    ctx.set("v", v)
    ctx.set("m", m)
    ctx.model_snow(1.4e5, 0.2, 10.0, 2.5e-2, 7.5e-3)
    ctx.set_colliders([plane])
    g = G.Grid(ctx, np.arange(n), h)
    g.compute_particle_volumes()
    gpu_its = []
    st = g.update_grid_velocities(dt, 200, 1e-5, debug=gpu_its.append)
    ref = {}
    for prec in ("float", "double"):
        pd = oracle.ParticleData.create(n, prec)
        pd.set("p", x)
        pd.set("v", v)
        pd.set("m", m)
        model = oracle.Model.snow(1.4e5, 0.2, 10.0, 2.5e-2, 7.5e-3, prec)
        model.set_particles(pd)
        og = oracle.Grid(pd, np.arange(n), float(np.float32(h)))
        og.compute_particle_volumes()
        cols = oracle.Colliders.create(prec)
        cols.add_plane(plane[0], v=plane[1], friction=plane[2], sticky=plane[3])
        fields = oracle.Fields.create(prec)
        ref[prec] = og.update_grid_velocities_iterates(dt, model, cols, fields, 200, 1e-5)
    i32, i64 = ref["float"], ref["double"]
    print("[gpu-case] iterates: gpu %d, fp32 oracle %d, fp64 oracle %d" % (len(gpu_its), len(i32), len(i64)),
          flush=True)
    assert len(gpu_its) == st["iterations"] > 3
    assert abs(len(gpu_its) - len(i32)) <= 1
    scale = np.abs(i64[-1]).max()
    k = min(len(gpu_its), len(i32), len(i64))
    worst = 0.0
    for j in range(k):
        e_gpu = np.abs(gpu_its[j] - i64[j]).max() / scale
        e_32 = np.abs(i32[j] - i64[j]).max() / scale
        worst = max(worst, e_gpu / max(3.0 * e_32, 1e-5))
        if j < 4 or j == k - 1:
            print("[gpu-case] iterate %d: gpu-fp64 %.2e, fp32-fp64 %.2e" % (j, e_gpu, e_32), flush=True)
    assert worst <= 1.0
