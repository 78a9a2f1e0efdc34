"""Config-4 geometry at test size: eight seeded boxes and spheres (12,000 particles, h = 1/48)
dropped at 1 m/s onto the ground plane and the analytic sphere collider (friction 0.2 on both),
stiff snow (E0 = 4.8e5 Pa, xi = 10, dt = 2e-3), 16 steps, against the FP64 oracle. Contact with
both supports and the plastic clamp are discontinuous, so the reference's own FP32 path departs
from FP64; the GPU stays inside that envelope (each quantile of its deviation from FP64 within 3x
the FP32 oracle's, plus a 1e-6 h floor for positions and 1e-6 for F), and so does the total
momentum. The analytic sphere's contact is checked directly: particles of the four inner objects
rest on it (none inside it by more than 0.5 h) and their vertical momentum stays in the same
envelope."""
import numpy as np
import pytest

import mpmsnow_amd as G
from mpmsnow_amd import scenes
from test_gpu_parity import F32, F64, say

pytestmark = pytest.mark.gpu

N, H_INV, STEPS = 12000, 48, 16


def oracle_sim(oracle, prec, sc):
    mat = sc["material"]
    cols = oracle.Colliders.create(prec)
    for c, vv, mu, st in sc["planes"]:
        cols.add_plane(c, v=vv, friction=mu, sticky=st)
    for c, r, vv, mu, st in sc["spheres"]:
        cols.add_sphere(c, r, v=vv, friction=mu, sticky=st)
    fields = oracle.Fields.create(prec)
    fields.add_gravity(sc["gravity"])
    model = oracle.Model.snow(mat["E"], mat["nu"], mat["xi"], mat["theta_c"], mat["theta_s"], prec)
    o = oracle.Sim(sc["x"], sc["m"], float(np.float32(sc["h"])), model, cols, fields, prec=prec)
    o.pd.set("v", sc["v"])
    o._keep = (cols, fields, model)
    return o


def test_stiff_drop_on_plane_and_sphere_within_fp32_envelope(oracle):
    say("stiff drop %d particles, h=1/%d, plane + analytic sphere, %d steps" % (N, H_INV, STEPS))
    sc = scenes.stiff_drop(n=N, seed=4, h=1.0 / H_INV)
    h, n, m = sc["h"], N, sc["m"]
    sim = G.Sim(sc["x"], m, h, material=sc["material"], planes=sc["planes"], spheres=sc["spheres"],
                gravity=[sc["gravity"]], v=sc["v"])
    osims = {prec: oracle_sim(oracle, prec, sc) for prec in (F64, F32)}
    iters = []
    for step in range(STEPS):
        st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
        iters.append(st["max_iterations"])
        for o in osims.values():
            o.advance(sc["dt"], sc["max_iters"], sc["tol"])
    say("GPU max CR iterations per step: %s" % iters)
    ref = osims[F64]
    v64 = ref.pd.get("v")
    mom = (m[:, None] * sim.get("v")).sum(0)
    mom64 = (m[:, None] * v64).sum(0)
    mom32 = (m[:, None] * osims[F32].pd.get("v")).sum(0)
    say("momentum gpu %s fp32 %s fp64 %s" % (mom, mom32, mom64))
    assert np.abs(mom - mom64).max() <= 3 * np.abs(mom32 - mom64).max() + 1e-5 * np.abs(mom64).max()
    for name, floor in (("p", 1e-6 * h), ("F", 1e-6)):
        r = ref.pd.get(name).reshape(n, -1)
        dg = np.abs(sim.get(name).reshape(n, -1) - r).max(1)
        d32 = np.abs(osims[F32].pd.get(name).reshape(n, -1) - r).max(1)
        for q in (50, 90, 99, 100):
            a, b = np.percentile(dg, q), np.percentile(d32, q)
            say("%s q%d: gpu %.3e  fp32 reference %.3e" % (name, q, a, b))
            assert a <= 3 * b + floor, (name, q, a, b)
    # the inner four objects rest on the analytic sphere
    (c, r, _, _, _), = sc["spheres"]
    x = sim.get("p")
    d = np.sqrt(((x - np.asarray(c)) ** 2).sum(1)) - r
    counts = np.cumsum([0] + [o["count"] for o in sc["objects"]])
    inner = np.concatenate([np.arange(counts[s], counts[s + 1]) for s in (1, 2, 5, 6)])
    say("inner objects: min distance to the sphere collider %.3e (h = %.4f)" % (d[inner].min(), h))
    assert d[inner].min() < h  # in contact
    assert d.min() > -0.5 * h  # not pushed through
    py, py64 = (m[inner] * sim.get("v")[inner, 1]).sum(), (m[inner] * v64[inner, 1]).sum()
    py32 = (m[inner] * osims[F32].pd.get("v")[inner, 1]).sum()
    say("inner objects' vertical momentum gpu %.6e fp32 %.6e fp64 %.6e" % (py, py32, py64))
    assert abs(py - py64) <= 3 * abs(py32 - py64) + 1e-4 * abs(py64)
