"""Body partition (Sim::calculateBodies, Sim.cpp:136-203) on the GPU against the CPU oracle's
KdTree flood fill: every body's particle list and the ballistic list must match exactly (index
work is bit-exact).

The clouds exercise the linking passes: one dense component holding most particles (the
sampled vote finds it and its particles skip the second pass), two equal halves and eight
separate spheres (all skipped: their boxes lie farther apart than h), many mid-size clusters
with no dominant one, chains whose spacing straddles h (pairs decided by the float distance
test), isolated particles only (no component), clusters at negative cell coordinates, a slab
spanning hundreds of the first pass's workgroups, a wide sparse domain whose cell blocks alias
in the bucket table, and large components whose boxes come within h of each other (0.5 h: one
body, which the first pass may leave in two parts and the second then joins from the part that
is not skipped; 1.0005 h: two bodies, inside the skip test's margin).
"""
import numpy as np
import pytest

import mpmsnow_amd as G
from mpmsnow_amd import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print("[bodies] " + msg, flush=True)


def oracle_bodies(oracle, x, h):
    prec = "float"
    model = oracle.Model.snow(1.4e5, 0.2, 10.0, 2.5e-2, 7.5e-3, prec)
    cols = oracle.Colliders.create(prec)
    fields = oracle.Fields.create(prec)
    sim = oracle.Sim(x, np.ones(len(x), np.float32), float(np.float32(h)), model, cols, fields, prec=prec)
    return sim.bodies(), sim.ballistic()


def gpu_bodies(x, h):
    sim = G.Sim(x, np.ones(len(x), np.float32), h, material=dict(scenes.SNOW))
    return sim.bodies(), sim.ballistic()


def clouds():
    rng = np.random.default_rng(42)
    h = 1.0 / 64
    out = {}
    big = scenes.sphere(3, 20000, (0.5, 0.5, 0.5), 0.12)
    specks = rng.uniform(0, 1, (300, 3)).astype(np.float32)
    out["one dense component + specks"] = np.concatenate([big, specks])
    a = scenes.sphere(4, 10000, (0.3, 0.5, 0.5), 0.1)
    b = scenes.sphere(5, 10000, (0.7, 0.5, 0.5), 0.1)
    out["two equal components"] = np.concatenate([a, b])
    parts = []
    for k in range(60):
        c = rng.uniform(0.1, 0.9, 3)
        parts.append(scenes.sphere(100 + k, int(rng.integers(2, 400)), c, float(rng.uniform(0.2, 3.0)) * h))
    out["60 clusters"] = np.concatenate(parts)
    chains = []
    for k in range(40):
        p0 = rng.uniform(0.1, 0.9, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        step = h * (1.0 + rng.uniform(-2e-6, 2e-6, 30))
        chains.append(p0 + np.cumsum(step)[:, None] * d)
    out["chains at spacing h"] = np.concatenate(chains).astype(np.float32)
    g = np.stack(np.meshgrid(*[np.arange(12)] * 3, indexing="ij"), -1).reshape(-1, 3)
    out["isolated only"] = (0.1 + 1.5 * h * g).astype(np.float32)
    # negative cell coordinates (the bucket of a cell takes its 4x4x4 block and the position in it)
    parts = []
    for k in range(300):
        c = rng.uniform(-0.3, 0.3, 3)
        parts.append(scenes.sphere(500 + k, int(rng.integers(2, 300)), c, float(rng.uniform(0.3, 2.0)) * h))
    out["300 clusters around the origin"] = np.concatenate(parts)
    # a slab of many 256-particle workgroups, so that pass 1 joins most links across workgroups
    slab = scenes.box(7, 120000, (0.2, 0.0, 0.3), (0.8, 0.1, 0.7))
    out["slab, 120k particles"] = np.concatenate([slab, rng.uniform(0, 1, (200, 3)).astype(np.float32)])
    # a wide sparse domain: the blocks alias in the bucket table
    wide = rng.uniform(-10.0, 10.0, (3000, 3)).astype(np.float32)
    out["sparse wide domain + sphere"] = np.concatenate([wide, scenes.sphere(8, 5000, (0.0, 0.0, 0.0), 0.05)])
    # eight large components far apart: the second pass skips all of them
    out["eight separate spheres"] = np.concatenate([
        scenes.sphere(900 + k, 4000, (0.3 + 0.35 * (k & 1), 0.3 + 0.35 * ((k >> 1) & 1), 0.3 + 0.35 * (k >> 2)), 0.1)
        for k in range(8)])
    # large components whose boxes come within the link distance: A and B 0.5 h apart (one body;
    # if pass 1 leaves them apart, only one of them is skipped and pass 2 over the other joins
    # them), B and C 3 h apart, A and C far apart (both may be skipped)
    a = scenes.box(910, 8000, (0.1, 0.4, 0.4), (0.3, 0.6, 0.6))
    b = scenes.box(911, 7000, (0.3 + 0.5 * h, 0.4, 0.4), (0.5, 0.6, 0.6))
    c = scenes.box(912, 6000, (0.5 + 3.0 * h, 0.4, 0.4), (0.7, 0.6, 0.6))
    out["three boxes, the middle one 0.5 h from the first"] = np.concatenate([a, b, c])
    # two large components a hair over h apart (1.0005 h: two bodies, inside the boxes' margin)
    a = scenes.box(913, 8000, (0.1, 0.4, 0.4), (0.3, 0.6, 0.6))
    b = scenes.box(914, 8000, (0.3 + 1.0005 * h, 0.4, 0.4), (0.5, 0.6, 0.6))
    out["two boxes 1.0005 h apart"] = np.concatenate([a, b])
    return out, h


def test_body_partition_matches_oracle(oracle):
    cases, h = clouds()
    for name, x in cases.items():
        say("%s: %d particles" % (name, len(x)))
        gb, gball = gpu_bodies(x, h)
        ob, oball = oracle_bodies(oracle, x, h)
        say("  bodies gpu %d oracle %d, ballistic gpu %d oracle %d, largest %d" % (
            len(gb), len(ob), len(gball), len(oball), max([len(b) for b in ob] + [0])))
        assert len(gb) == len(ob), name
        for k, (p, q) in enumerate(zip(gb, ob)):
            assert np.array_equal(np.sort(p), np.sort(q)), (name, k)
        assert np.array_equal(np.sort(gball), np.sort(oball)), name
