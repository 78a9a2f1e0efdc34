"""The full-FP64 GPU step (mpm_sim64_*) against the FP64 oracle stepping the same seeded scenes.

Both run the reference's algorithm in double (Sim.cpp:57-134); they differ only in summation order
(FP64 atomics against the oracle's serial loops), in fused multiply-adds and in the SVD's
rounding, all at about 1e-16 relative. The bars below are six or more orders of magnitude under
what FP32 arithmetic gives on the same steps (1e-6 to 1e-4): per step the body partition, every
body's CR iteration count and residual history, and after the last step positions, velocities and
F of every particle."""
import numpy as np
import pytest

import mpmsnow_amd as G
from mpmsnow_amd import scenes

pytestmark = pytest.mark.gpu


def oracle64(O, sc):
    mat = sc["material"]
    cols = O.Colliders.create("double")
    for c, v, mu, st in sc["planes"]:
        cols.add_plane(c, v=v, friction=mu, sticky=st)
    for c, r, v, mu, st in sc.get("spheres", ()):
        cols.add_sphere(c, r, v=v, friction=mu, sticky=st)
    f = O.Fields.create("double")
    f.add_gravity(sc["gravity"])
    model = O.Model.snow(mat["E"], mat["nu"], mat["xi"], mat["theta_c"], mat["theta_s"], "double")
    extra = dict(mu0=sc["mu0"], lambda0=sc["lambda0"], xi=sc["xi"]) if "mu0" in sc else {}
    sim = O.Sim(sc["x"], sc["m"], float(np.float32(sc["h"])), model, cols, f, prec="double", **extra)
    sim.pd.set("v", sc["v"])
    sim._keep = (cols, f, model)
    return sim


def gpu64(sc):
    extra = dict(mu0=sc["mu0"], lambda0=sc["lambda0"], xi=sc["xi"]) if "mu0" in sc else {}
    return G.Sim64(sc["x"], sc["m"], float(np.float32(sc["h"])), sc["material"], planes=sc["planes"],
                   spheres=sc.get("spheres", ()), gravity=[sc["gravity"]], v=sc["v"], **extra)


def _hist_dev(sa, sb):
    """Largest relative difference of two runs' residual histories over a step's bodies: relative to
    the entry, or to 1e-16 |b|^2 (|r| a thousandth of the tolerance) for entries at the rounding
    floor (a free-fall solve ends after one iteration with |r|^2 ~ 1e-24 |b|^2)."""
    worst = 0.0
    for ref, got in zip(sa, sb):
        L = min(len(ref["residual_norm2"]), len(got["residual_norm2"]))
        ho, hg = ref["residual_norm2"][:L], got["residual_norm2"][:L]
        if L:
            worst = max(worst, float(np.max(np.abs(hg - ho) / np.maximum(ho, 1e-16 * ref["bnorm2"]))))
    return worst


def _state_dev(a, b, h):
    dx = float(np.abs(a.get("p") - b.get("p")).max()) / h
    dF = float(np.abs(a.get("F") - b.get("F")).max())
    return dx, dF


class _OracleView:
    def __init__(self, o):
        self.o = o

    def get(self, name):
        return self.o.pd.get(name)


def run_pair(O, sc, steps, name):
    """Steps the GPU's FP64 Sim and the FP64 oracle together. Every step: the same bodies, particle
    counts and CR iteration counts. Returns per step (GPU vs oracle) the position and F deviation
    and the residual-history deviation, and the final velocity deviation."""
    h = float(np.float32(sc["h"]))
    g = gpu64(sc)
    o = oracle64(O, sc)
    rows = []
    for step in range(steps):
        st = g.advance(sc["dt"], sc["max_iters"], sc["tol"])
        o.advance(sc["dt"], sc["max_iters"], sc["tol"])
        ost = o.last_stats()
        assert st["num_solved"] == len(ost), (name, step, st["num_solved"], len(ost))
        assert g.num_bodies() == len(o.bodies()), (name, step)
        gst = [g.body_stats(b) for b in range(len(ost))]
        for b, (ref, got) in enumerate(zip(ost, gst)):
            assert got["particles"] == ref["particles"], (name, step, b)
            assert got["iterations"] == ref["iterations"], (name, step, b, got["iterations"], ref["iterations"])
        row = {"gpu": _state_dev(g, _OracleView(o), h) + (_hist_dev(ost, gst),)}
        rows.append(row)
        print("[fp64] %s step %d: GPU vs FP64 oracle |dx| %.2e h |dF| %.2e history %.2e" % (
            name, step + 1, *row["gpu"]), flush=True)
    dv = float(np.abs(g.get("v") - o.pd.get("v")).max() / max(np.abs(o.pd.get("v")).max(), 1e-30))
    return rows, dv


def test_fp64_snowball_contact_matches_fp64_oracle(oracle):
    print("[gpu-case] fp64 step: snowball hitting the ground, 2000 particles, 12 steps", flush=True)
    sc = scenes.snowball_drop(n=2000, seed=11, contact=True)
    rows, dv = run_pair(oracle, sc, 12, "snowball contact")
    for r in rows:
        dx, dF, hist = r["gpu"]
        assert dx < 1e-9 and dF < 1e-9 and hist < 1e-7, r
    assert dv < 1e-9


def test_fp64_two_snowballs_match_fp64_oracle(oracle):
    """Two snowballs meeting, through the first five steps of contact (the solves run 31-51
    iterations from step 3). At step 8 the scene's map is discontinuous: two runs of the FP64
    oracle 1e-15 apart at the start, equal to 1.2e-12 h after step 7, differ by 2.3e-8 h (all
    particles, through the implicit solve) after step 8 and 1.4e-3 in F after step 12; which
    particles switch branch, and so by how much two runs part, depends on the last bits. Steps 1-7
    are compared in full: equal bodies and CR iteration counts, positions and F within 1e-11,
    residual histories within 1e-9 relative."""
    # the two-snowball scene's particle density (500k per 0.15 m ball) at 1500 per ball, closing at
    # 20 m/s so that they meet within the run
    print("[gpu-case] fp64 step: two snowballs colliding, 3000 particles, 7 steps", flush=True)
    sc = scenes.two_snowballs(n=3000, seed=5, r=0.0216, speed=10.0)
    rows, dv = run_pair(oracle, sc, 7, "two snowballs")
    for r in rows:
        dx, dF, hist = r["gpu"]
        assert dx < 1e-11 and dF < 1e-11 and hist < 1e-9, r
    assert dv < 1e-11


def test_fp64_slope_slab_matches_fp64_oracle(oracle):
    """Per-particle E0 and xi, tilted gravity over a frictional plane. From the reference's initial
    state (Ginv = I where F = I gives I/2, SnowConstitutiveModel.cpp:123-169) the first solve is
    ill-conditioned and runs to the iteration cap; there every iteration amplifies rounding
    (GPU vs oracle, measured: 2e-15 h after 20 iterations, 2e-11 h after 50, 1e-7 h after 100,
    1e-6 h after 1000), so that step is compared at a 20-iteration cap. From the consistent state
    Ginv = I/2 the first solve converges (40 iterations) and is compared in full."""
    import precision_study as PS

    print("[gpu-case] fp64 step: slope slab, 4000 particles, first step capped at 20 iterations", flush=True)
    sc, _ = PS.study_scene(4000)
    sc = dict(sc, max_iters=20)
    rows, dv = run_pair(oracle, sc, 1, "slope slab, 20-iteration cap")
    dx, dF, hist = rows[0]["gpu"]
    assert dx < 1e-12 and dF < 1e-12 and hist < 1e-9 and dv < 1e-12, rows
    print("[gpu-case] fp64 step: slope slab, 4000 particles, one step from Ginv = I/2", flush=True)
    sc, _ = PS.study_scene(4000)
    n = len(sc["m"])
    g, o = gpu64(sc), oracle64(oracle, sc)
    half = np.tile(0.5 * np.eye(3), (n, 1, 1))
    g.set("Ginv", half)
    o.pd.set("Ginv", half)
    g.advance(sc["dt"], sc["max_iters"], sc["tol"])
    o.advance(sc["dt"], sc["max_iters"], sc["tol"])
    ost = o.last_stats()
    gst = [g.body_stats(b) for b in range(len(ost))]
    assert [r["iterations"] for r in gst] == [r["iterations"] for r in ost]
    dx, dF = _state_dev(g, _OracleView(o), float(np.float32(sc["h"])))
    hist = _hist_dev(ost, gst)
    print("[fp64] slope slab from Ginv = I/2: %d iterations, |dx| %.2e h |dF| %.2e history %.2e" % (
        ost[0]["iterations"], dx, dF, hist), flush=True)
    assert dx < 1e-12 and dF < 1e-12 and hist < 1e-7


def test_fp32_step_100_contact_steps_against_the_fp64_step():
    """The FP32-bulk step (mpm_sim_advance) against the FP64 step over 100 steps of a snowball
    (the snowball-drop scene: 20k particles, r = 0.1 m, h = 1/64, dt = 1e-3) driven into the frictional ground,
    so that the solves run in contact from step 10 on. Positions: 99th percentile within 1e-4 h
    (measured 3.3e-5 h), median within 2e-5 h, maximum within 5e-4 h (measured 1.5e-4 h, a few
    particles at the contact's edge); F: 99th percentile within 1e-4 (measured 2.4e-5)."""
    print("[gpu-case] fp32 vs fp64 step: snowball contact 20k particles, 100 steps", flush=True)
    sc = scenes.snowball_drop(n=20000, seed=3, contact=True)
    h = float(np.float32(sc["h"]))
    g32 = G.Sim(sc["x"], sc["m"], h, material=sc["material"], planes=sc["planes"], gravity=[sc["gravity"]],
                v=sc["v"])
    g64 = gpu64(sc)
    for _ in range(100):
        g32.advance(sc["dt"], sc["max_iters"], sc["tol"])
        g64.advance(sc["dt"], sc["max_iters"], sc["tol"])
    dx = np.abs(g32.get("p") - g64.get("p")).max(1) / h
    dF = np.abs(g32.get("F").reshape(-1, 9) - g64.get("F").reshape(-1, 9)).max(1)
    print("[fp64] 100 contact steps, FP32 bulk vs FP64: x max %.2e p99 %.2e median %.2e h; F max %.2e p99 %.2e" % (
        dx.max(), np.percentile(dx, 99), np.median(dx), dF.max(), np.percentile(dF, 99)), flush=True)
    assert np.percentile(dx, 99) < 1e-4 and np.median(dx) < 2e-5 and dx.max() < 5e-4
    assert np.percentile(dF, 99) < 1e-4
