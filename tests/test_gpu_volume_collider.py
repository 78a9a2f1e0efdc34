"""GPU parity with a level-set volume collider (the reference's VDBCollisionObject, houdiniPlugin/
VDBCollisionObject.cpp:11-37, as a dense volume): explicit momenta, collision flags, collision
velocities of a sampled velocity field and the implicit solve against the oracle; a ground plane
given as a volume against the analytic plane over a contact run."""
import numpy as np
import pytest

import mpmsnow_amd as G
from test_gpu_parity import F32, F64, REL, Pair, implicit_pair, relerr, say

pytestmark = pytest.mark.gpu


def sphere_volume():
    c, r = np.array([0.6, -0.9, 0.7]), 1.1
    lo, voxel, dims = np.array([-2.0, -2.5, -2.0]), 0.125, (45, 41, 45)
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    P = np.stack([lo[0] + i * voxel, lo[1] + j * voxel, lo[2] + k * voxel], -1)
    phi = (np.linalg.norm(P - c, axis=-1) - r).astype(np.float32)
    # a rotating velocity field about the sphere's centre plus a drift
    w = np.array([0.3, -0.2, 0.5])
    vel = (np.cross(w, P - c) + np.array([0.05, 0.1, -0.02])).astype(np.float32)
    return phi, lo, voxel, vel


def pair_with_volume(oracle):
    P, h, _ = implicit_pair(oracle)
    phi, lo, voxel, vel = sphere_volume()
    planes = [((-1, 0, 0, 1.5), (-0.2, -0.1, -0.4), 0.5, False)]
    P.ctx.set_colliders(planes, volumes=[(phi, lo, voxel, vel, 0.4, False)])
    cols = {}
    for prec in (F32, F64):
        cs = oracle.Colliders.create(prec)
        for c, v, mu, st in planes:
            cs.add_plane(c, v=v, friction=mu, sticky=st)
        cs.add_volume(phi.astype(np.float64), lo, voxel, v=vel.astype(np.float64), friction=0.4)
        cols[prec] = cs
    return P, h, cols


def test_volume_collider_explicit_momenta_and_collisions(oracle):
    say("explicit_momenta plane + level-set sphere with a velocity field")
    P, h, cols = pair_with_volume(oracle)
    inds = np.arange(P.n)
    g = G.Grid(P.ctx, inds, h)
    g.compute_particle_volumes()
    em, col = g.explicit_momenta(0.002)
    vc = g.collision_velocities(col)
    n_hit = int((col >= 0).sum())
    say("nodes in contact: %d (volume: %d)" % (n_hit, int((col == 1).sum())))
    assert (col == 1).sum() >= 5
    for prec, tol in ((F32, 1e-6), (F64, REL)):
        og = oracle.Grid(P.o[prec], inds, h if prec == F32 else float(np.float32(h)))
        og.compute_particle_volumes()
        fields = oracle.Fields.create(prec)
        emo, colo = og.explicit_momenta(0.002, P.om[prec], cols[prec], fields)
        if prec == F32:
            assert (col == colo).all()  # the same float arithmetic decides every flag
        else:
            assert (col == colo).mean() > 0.995
        k = np.repeat(col == colo, 3)
        assert relerr(em[k], emo[k]) < tol
        vco = og.collision_velocities(cols[prec], colo)
        assert np.abs(vc[k] - vco[k]).max() < 1e-6


def test_volume_collider_implicit_solve(oracle):
    say("implicit_solve plane + level-set sphere maxIters=40 tol=0")
    P, h, cols = pair_with_volume(oracle)
    inds = np.arange(P.n)
    g = G.Grid(P.ctx, inds, h)
    og = oracle.Grid(P.o[F64], inds, h)
    g.compute_particle_volumes()
    og.compute_particle_volumes()
    fields = oracle.Fields.create(F64)
    st = g.update_grid_velocities(0.002, 40, 0.0)
    sto = og.update_grid_velocities(0.002, P.om[F64], cols[F64], fields, 40, 0.0)
    assert st["iterations"] == sto["iterations"]
    assert relerr(g.get("velocities"), og.get("velocities")) < 1e-4


def test_ground_volume_matches_ground_plane():
    """A ball dropped on y = 0 given once as the analytic plane and once as a plane volume."""
    say("contact run: ground plane vs ground volume, 40 steps")
    rng = np.random.default_rng(4)
    n = 4000
    p = rng.uniform(-1, 1, (3 * n, 3))
    x = (np.array([0.5, 0.14, 0.5]) + 0.1 * p[(p * p).sum(1) < 1][:n]).astype(np.float32)
    m = np.full(n, 1e-4, np.float32)
    v = np.tile(np.float32([0.3, -2.0, 0.0]), (n, 1))
    mat = dict(E=1.4e5, nu=0.2, xi=10.0, theta_c=2.5e-2, theta_s=7.5e-3)
    h = 1.0 / 64
    lo, voxel, dims = (-0.5, -0.5, -0.5), 0.0625, (33, 33, 33)
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    phi = (lo[1] + j * voxel).astype(np.float32)  # phi = y
    runs = []
    for use_volume in (False, True):
        if use_volume:
            sim = G.Sim(x, m, h, material=mat, volumes=[(phi, lo, voxel, (0, 0, 0), 0.2, False)],
                        gravity=[(0, -9.81, 0)], v=v)
        else:
            sim = G.Sim(x, m, h, material=mat, planes=[((0, 1, 0, 0), (0, 0, 0), 0.2, False)],
                        gravity=[(0, -9.81, 0)], v=v)
        for _ in range(40):
            sim.advance(1e-3, 200, 1e-5)
        runs.append(sim.get("p"))
    assert runs[0].min(0)[1] > -2 * h  # the ball met the ground and stayed above it
    d = np.abs(runs[0] - runs[1]).max()
    say("max position difference %.3g (h = %.3g)" % (d, h))
    assert d < 1e-3 * h


def test_rejected_collider_set_keeps_the_previous_one(oracle):
    """mpm_set_colliders validates the whole set before it touches device memory: a call whose
    second volume is malformed (voxel size 0) fails and leaves the previous set, and the volume
    buffers its device copy points into, as they were, although the rejected set's first volume
    is larger than the buffers the previous one filled."""
    say("rejected collider set")
    P, h, cols = pair_with_volume(oracle)
    inds = np.arange(P.n)
    g = G.Grid(P.ctx, inds, h)
    g.compute_particle_volumes()
    em, col = g.explicit_momenta(0.002)
    phi, lo, voxel, vel = sphere_volume()
    big = np.ascontiguousarray(np.tile(phi, (3, 1, 1)))
    bigv = np.ascontiguousarray(np.tile(vel, (3, 1, 1, 1)))
    with pytest.raises(G.MpmError):
        P.ctx.set_colliders([((-1, 0, 0, 1.5), (0.0, 0.0, 0.0), 0.5, False)],
                            volumes=[(big, lo, voxel, bigv, 0.4, False), (phi, lo, 0.0, vel, 0.4, False)])
    g2 = G.Grid(P.ctx, inds, h)
    g2.compute_particle_volumes()
    em2, col2 = g2.explicit_momenta(0.002)
    assert (col2 == col).all()
    assert np.array_equal(em2, em)
