"""Level-set volume collider (the reference's VDBCollisionObject, houdiniPlugin/VDBCollisionObject.cpp:
11-37, re-expressed as a dense volume; the HDK sampling it calls is absent, so the sampling rule is
this build's own and is pinned here by what the mathematics fixes): trilinear interpolation
reproduces linear fields, so a plane given as a volume collides exactly like the analytic plane
(CollisionPlane.cpp:23-46) and a linear velocity field is sampled exactly; a sphere's sampled
distance stays within the interpolation error of the analytic sphere; grid points return their
own values; positions outside the volume clamp to its boundary."""
import numpy as np
import pytest


def plane_volume(n, d, lo, voxel, dims):
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    x = lo[0] + i * voxel
    y = lo[1] + j * voxel
    z = lo[2] + k * voxel
    return n[0] * x + n[1] * y + n[2] * z + d


@pytest.mark.parametrize("prec", ["float", "double"])
def test_plane_volume_collides_like_the_plane(oracle, prec):
    rng = np.random.default_rng(1)
    n = np.array([0.3, 0.9, -0.2]); n /= np.linalg.norm(n)
    d = -0.25
    lo, voxel, dims = (0.0, 0.0, 0.0), 0.0625, (17, 17, 17)  # exact binary grid
    phi = plane_volume(n, d, lo, voxel, dims)
    vol = oracle.Colliders.create(prec)
    vol.add_volume(phi, lo, voxel, v=(0.1, -0.2, 0.05), friction=0.4)
    pl = oracle.Colliders.create(prec)
    pl.add_plane((n[0], n[1], n[2], d), v=(0.1, -0.2, 0.05), friction=0.4)
    tol = 2e-5 if prec == "float" else 1e-12
    checked = 0
    for _ in range(400):
        x = rng.uniform(0.1, 0.9, 3)
        if abs(n @ x + d) < 1e-3:
            continue
        v = rng.normal(size=3)
        fv = rng.normal(scale=0.1, size=3)
        f1, v1 = vol.collide(v, x, fv)
        f2, v2 = pl.collide(v, x, fv)
        assert f1 == f2
        assert np.abs(v1 - v2).max() <= tol * max(1.0, np.abs(v2).max())
        checked += 1
    assert checked > 300


def test_volume_sampling_exact_cases(oracle):
    rng = np.random.default_rng(2)
    lo, voxel, dims = (-0.5, 0.0, 0.25), 0.125, (9, 7, 5)
    # grid points return their own values. phi <= 0 is read through a second collider that always
    # intersects (phi = -1): the set reports -2 (two objects) exactly when the volume's phi <= 0
    phi = rng.uniform(-1, 1, (dims[2], dims[1], dims[0]))
    phi[phi == 0] = 0.5
    vol = oracle.Colliders.create("double")
    vol.add_volume(phi, lo, voxel, v=(0, 0, 0), friction=0.0)
    vol.add_plane((0, 0, 0, -1))
    for _ in range(200):
        i, j, k = rng.integers(0, dims[0]), rng.integers(0, dims[1]), rng.integers(0, dims[2])
        x = np.array([lo[0] + i * voxel, lo[1] + j * voxel, lo[2] + k * voxel])
        flag, _ = vol.collide(np.zeros(3), x)
        assert (flag == -2) == (phi[k, j, i] <= 0)
        # outside the volume the sample clamps to the nearest boundary point
        xo = x + np.array([(i == dims[0] - 1) * 3.0 - (i == 0) * 3.0, 0, 0])
        flag_o, _ = vol.collide(np.zeros(3), xo)
        assert (flag_o == -2) == (phi[k, j, i] <= 0)
    # a linear velocity field is reproduced exactly (trilinear interpolation of an affine map)
    A = rng.normal(size=(3, 3)); b = rng.normal(size=3)
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    P = np.stack([lo[0] + i * voxel, lo[1] + j * voxel, lo[2] + k * voxel], -1)
    vel = P @ A.T + b
    inside = np.full((dims[2], dims[1], dims[0]), -1.0)  # everything collides
    vv = oracle.Colliders.create("double")
    vv.add_volume(inside, lo, voxel, v=vel, friction=0.0, sticky=True)
    for _ in range(200):
        x = np.array([rng.uniform(lo[0], lo[0] + 8 * voxel), rng.uniform(0, 6 * voxel),
                      rng.uniform(lo[2], lo[2] + 4 * voxel)])
        # sticky: v becomes 0 relative to the object, then the object's velocity is added back
        flag, v = vv.collide(rng.normal(size=3), x, add_collision_velocity=True)
        if flag >= 0:
            assert np.allclose(v, A @ x + b, rtol=0, atol=1e-12)


def test_sphere_volume_close_to_the_sphere(oracle):
    c, r = np.array([0.5, 0.5, 0.5]), 0.2
    lo, voxel, dims = (0.2, 0.2, 0.2), 0.0125, (49, 49, 49)
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    P = np.stack([lo[0] + i * voxel, lo[1] + j * voxel, lo[2] + k * voxel], -1)
    phi = np.linalg.norm(P - c, axis=-1) - r
    vol = oracle.Colliders.create("double")
    vol.add_volume(phi, lo, voxel, friction=0.0)
    sph = oracle.Colliders.create("double")
    sph.add_sphere(c, r, friction=0.0)
    rng = np.random.default_rng(3)
    agree = 0
    total = 0
    for _ in range(500):
        x = c + rng.uniform(-0.25, 0.25, 3)
        d = np.linalg.norm(x - c) - r
        if abs(d) < 2 * voxel * voxel / r:  # within the interpolation error of the surface
            continue
        v = rng.normal(size=3)
        f1, v1 = vol.collide(v, x)
        f2, v2 = sph.collide(v, x)
        total += 1
        if f1 == f2:
            agree += 1
            # normals differ by the central difference error, O(voxel^2 / r^2)
            assert np.abs(v1 - v2).max() <= 1e-2 * max(1.0, np.abs(v2).max())
    assert agree == total


def test_nan_position_samples_inside_the_volume(oracle):
    """A NaN position reads the volume's first point (never outside the array): with phi > 0
    there, no collision."""
    phi = np.ones((3, 3, 3))
    for prec in ("float", "double"):
        vol = oracle.Colliders.create(prec)
        vol.add_volume(phi, (0, 0, 0), 0.5)
        flag, _ = vol.collide(np.array([0.0, -1.0, 0.0]), np.array([np.nan, 0.2, np.nan]))
        assert flag == -1
