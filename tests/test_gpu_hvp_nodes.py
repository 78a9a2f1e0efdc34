"""The HVP's error node by node (force differentials, Grid.cpp:549-632) on a slope-slab state, against
the FP64 oracle.

The GPU scatters each chunk's contributions in 32-bit fixed point with one scale per component,
set by the chunk's largest contribution (resolution 2^-30 of it). The error is therefore
absolute at that resolution: relative to its own value a node is exact to FP32 rounding when
its value is a sizeable fraction of the chunk's largest, and carries up to the whole quantum
when only far stencil-edge contributions reach it (nodes outside a body's surface, of negligible
mass). The test holds the whole vector to 1e-5 relative and every node to 1e-4 of its own value
plus 1e-5 of the vector's largest, and prints the error by magnitude band."""
import numpy as np
import pytest

import mpmsnow_amd as G

pytestmark = pytest.mark.gpu


def test_hvp_error_by_node_magnitude(oracle):
    import precision_study as PS

    print("[gpu-case] hvp per-node error: slope slab 30000 particles after 6 oracle steps", flush=True)
    sc, _ = PS.study_scene(30000)
    a = PS._oracle(oracle, sc, "double")
    for _ in range(6):
        a.advance(sc["dt"], sc["max_iters"], sc["tol"])
    keys = ("p", "v", "F", "volume", "Fp", "FinvTrans", "R", "Ginv", "J", "mu", "lambda")
    state = {k: a.pd.get(k) for k in keys}
    b0 = max(a.bodies(), key=len)
    h = float(np.float32(sc["h"]))
    model = a._keep[2]
    og = oracle.Grid(a.pd, b0, h)
    ctx = G.Context(0)
    ctx.particles_create(len(sc["m"]))
    ctx.set("p", state["p"])
    ctx.set("m", sc["m"])
    mat = sc["material"]
    ctx.model_snow(mat["E"], mat["nu"], mat["xi"], mat["theta_c"], mat["theta_s"])
    for k in ("mu0", "lambda0", "xi"):
        ctx.set(k, sc[k])
    for k in keys[1:]:
        ctx.set(k, state[k])
    ctx.set_colliders(sc["planes"])
    ctx.set_gravity([sc["gravity"]])
    g = G.Grid(ctx, b0, h)
    N = og.num_nodes
    mass = og.get("masses")
    dx = np.random.default_rng(1).normal(size=3 * N) * np.repeat(mass > 0, 3)
    d64 = og.force_differentials(model, dx)
    d32 = g.force_differentials(dx.astype(np.float32)).astype(np.float64)
    err = np.abs(d32 - d64)
    big = np.abs(d64).max()
    norm = np.linalg.norm(d32 - d64) / np.linalg.norm(d64)
    mag = np.abs(d64).reshape(N, 3).max(1)
    e = err.reshape(N, 3).max(1)
    on = mag > 0
    rel, frac = e[on] / mag[on], mag[on] / big
    for lo, hi in ((1e-1, 2.0), (1e-3, 1e-1), (1e-5, 1e-3), (0.0, 1e-5)):
        sel = (frac >= lo) & (frac < hi)
        if sel.any():
            print("[hvp] nodes with |df| in [%.0e, %.0e) of the largest: %5d, relative error median %.2e, max %.2e" % (
                lo, hi, sel.sum(), np.median(rel[sel]), rel[sel].max()), flush=True)
    print("[hvp] whole vector: relative norm error %.2e" % norm, flush=True)
    assert norm < 1e-5
    assert np.all(err <= 1e-4 * np.abs(d64) + 1e-5 * big)
