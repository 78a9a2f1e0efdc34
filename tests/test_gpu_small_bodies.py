"""The one-workgroup solver of small bodies against the batched CR loop, and the
bitwise reproducibility of a step.

A snowball (about 350 particles per 4x4x4-node tile, so tiles split into
several chunks) plus clusters of 2-600 particles (separate bodies, each within
the small-body limit of 8 node blocks) is advanced with the small-body solver
on and off. Both runs must give bitwise the same particle state and the same
per-body iteration counts, final residuals and residual histories: the solver
shares the batched kernels' node arithmetic, HVP scatter, reduction order and
scalar updates. A second run with the solver on must repeat the first bit for
bit (tiles are ordered by particle slot before they are split into chunks).
"""
import numpy as np
import pytest

import mpmsnow_amd as G
from mpmsnow_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print("[small] " + msg, flush=True)


def cluster_scene(seed=21):
    h = 1.0 / 64
    rng = np.random.default_rng(seed)
    parts = [scenes.sphere(seed, 6000, (0.5, 0.3, 0.5), 0.1)]
    vels = [np.tile(np.array([0.0, -1.0, 0.0], np.float32), (6000, 1))]
    sizes = [2, 3, 2, 5, 4, 2, 3, 40, 150, 600, 2, 3, 7, 2, 90, 2]
    k = 0
    for ix in range(4):
        for iz in range(4):
            n = sizes[k]
            k += 1
            c = (0.3 + 0.13 * ix, 0.62, 0.3 + 0.13 * iz)
            r = (0.35 if n < 10 else 1.2) * h
            parts.append(scenes.sphere(seed + 100 + k, n, c, r))
            v = rng.normal(scale=0.5, size=3).astype(np.float32)
            vels.append(np.tile(v, (n, 1)) + rng.normal(scale=0.05, size=(n, 3)).astype(np.float32))
    x = np.concatenate(parts).astype(np.float32)
    v = np.concatenate(vels).astype(np.float32)
    m = np.full(len(x), 400 * (4 / 3 * np.pi * 0.1 ** 3) / 6000, np.float32)
    return x, v, m, h


def run(limit, steps=4):
    x, v, m, h = cluster_scene()
    planes = [((0, 1, 0, 0), (0, 0, 0), 0.2, False)]
    sim = G.Sim(x, m, h, material=dict(scenes.SNOW), planes=planes, gravity=[(0.0, -9.81, 0.0)], v=v)
    sim.ctx.small_bodies(limit)
    sim.profile(["cr_small"])
    stats = []
    for _ in range(steps):
        st = sim.advance(1e-3, 1000, 1e-5)
        stats.append([sim.body_stats(b) for b in range(st["num_solved"])])
    launches = sim.profile_query()["cr_small"]["launches"]
    return sim, stats, launches


def same_runs(sa, sb):
    for step, (ra, rb) in enumerate(zip(sa, sb)):
        assert len(ra) == len(rb)
        its = []
        for x, y in zip(ra, rb):
            assert x["particles"] == y["particles"]
            assert x["iterations"] == y["iterations"], (step, x, y)
            assert x["early_exit"] == y["early_exit"]
            assert x["residual_norm2_last"] == y["residual_norm2_last"]
            assert np.array_equal(x["residual_norm2"], y["residual_norm2"])
            its.append((x["particles"], x["iterations"]))
        say("step %d (particles, iterations): %s" % (step, its))


def test_small_body_solver_is_bitwise_the_batched_loop():
    say("batched loop only")
    a, sa, la = run(0)
    say("small-body solver on")
    b, sb, lb = run(2048)
    assert la == 0 and lb >= 1, (la, lb)
    nb = [len(s) for s in sb]
    say("bodies per step %s; small-solver launches %d" % (nb, lb))
    assert min(nb) >= 10
    same_runs(sa, sb)
    for name in ("p", "v", "F"):
        assert np.array_equal(a.get(name), b.get(name)), name
    say("small-body solver on, again")
    c, sc, _ = run(2048)
    same_runs(sb, sc)
    for name in ("p", "v", "F"):
        assert np.array_equal(b.get(name), c.get(name)), name


def test_fragmenting_collision_small_solver_bitwise():
    """Two colliding snowballs (100k particles) shatter into fragments, many of them small bodies
    with several chunks; 90 steps with the small-body solver on and off must agree bit for bit
    (this caught a race between consecutive chunks of one workgroup's HVP)."""
    def run_collision(limit):
        sc = scenes.two_snowballs(n=100000, seed=3, r=0.07, speed=3.0)
        sim = G.Sim(sc["x"], sc["m"], sc["h"], material=sc["material"], planes=sc["planes"],
                    gravity=[sc["gravity"]], v=sc["v"])
        sim.ctx.small_bodies(limit)
        its = []
        for _ in range(90):
            st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
            its.append((st["num_bodies"], st["total_iterations"]))
        return sim, its
    say("collision 100k, small-body solver off")
    a, ia = run_collision(0)
    say("collision 100k, small-body solver on")
    b, ib = run_collision(2048)
    say("bodies / CR iterations at steps 30, 60, 89: %s" % [ib[k] for k in (30, 60, 89)])
    assert ia == ib
    for name in ("p", "v", "F"):
        assert np.array_equal(a.get(name), b.get(name)), name


def test_learned_handoff_small_solver_bitwise():
    """Learned mode with a network that every body abandons within a few iterations: at the
    handoff to the CR recurrences the small bodies still iterating go to the one-workgroup solver,
    whose restarted solves must be bitwise those of the batched loop (solver off)."""
    from oracle import lpc_oracle as L
    from test_gpu_lpc import random_params

    p = random_params(13)
    offs, _ = L.layer_offsets()
    w, b = offs[5]
    p[w:w + 27 * 16 * 16] *= 50.0

    def run_learned(limit):
        x, v, m, h = cluster_scene()
        planes = [((0, 1, 0, 0), (0, 0, 0), 0.2, False)]
        sim = G.Sim(x, m, h, material=dict(scenes.SNOW), planes=planes, gravity=[(0.0, -9.81, 0.0)], v=v)
        sim.ctx.small_bodies(limit)
        sim.ctx.lpc_set_params(p)
        sim.ctx.lpc_mode(1)
        sim.profile(["cr_small"])
        stats = []
        for _ in range(3):
            st = sim.advance(1e-3, 1000, 1e-5)
            stats.append([sim.body_stats(b) for b in range(st["num_solved"])])
        return sim, stats, sim.profile_query()["cr_small"]["launches"]

    say("clusters, learned mode with a bad network, small-body solver off / on")
    a, sa, la = run_learned(0)
    b, sb, lb = run_learned(2048)
    say("small-solver launches off %d, on %d; bodies per step %s" % (la, lb, [len(s) for s in sb]))
    assert la == 0 and lb > 0
    for ra, rb in zip(sa, sb):
        assert [(r["iterations"], r["fallbacks"]) for r in ra] == [(r["iterations"], r["fallbacks"]) for r in rb]
        for x, y in zip(ra, rb):
            assert np.array_equal(x["residual_norm2"], y["residual_norm2"])
    for name in ("p", "v", "F"):
        assert np.array_equal(a.get(name), b.get(name)), name
