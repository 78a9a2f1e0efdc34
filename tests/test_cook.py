"""The SOP cook driver's seeding (mpmsnow_amd/cpp/SnowCook.hpp, after the reference's
houdiniPlugin/SOP_MPMSim.cpp:167-196) on the CPU: the test program seeds the snowball level set
of tests/cpp/test_cook.cpp and the result must equal this closed-form restatement, particle for
particle in order: cell centres lo + (i + 1/2) h of a lattice of spacing h = gridSize / 2 over the
level set's bounding box (i < ceil(size / h) per axis, i outermost), kept where the trilinearly
sampled phi is negative."""
import os
import subprocess

import numpy as np

from test_cpp_layer import load_recipe

f32 = np.float32


def snowball():
    """The lattice of tests/cpp/test_cook.cpp's snowball(), in the same float arithmetic."""
    d = (20, 20, 20)
    vox = f32(1.0) / f32(64.0)
    o = np.array([0.35, 0.15, 0.35], dtype=f32)
    k, j, i = np.meshgrid(np.arange(d[2]), np.arange(d[1]), np.arange(d[0]), indexing="ij")
    x = o[0] + vox * i.astype(f32)
    y = o[1] + vox * j.astype(f32)
    z = o[2] + vox * k.astype(f32)
    dx, dy, dz = x - f32(0.5), y - f32(0.3), z - f32(0.5)
    phi = np.sqrt((dx * dx + dy * dy) + dz * dz) - f32(0.1)
    return phi.astype(f32), o, vox, d


def sample(phi, o, vox, d, p):
    """LevelSetCollisionObject::sample (trilinear, clamped; lerp a + t (b - a))."""
    i0, t = [], []
    for a in range(3):
        u = (p[a] - o[a]) / vox
        hi = f32(d[a] - 1)
        u = f32(0.0) if not u >= 0 else u
        u = hi if u > hi else u
        ii = min(int(np.floor(u)), d[a] - 2)
        i0.append(ii)
        t.append(f32(u - f32(ii)))

    def lerp(a, b, tt):
        return f32(a + f32(tt * f32(b - a)))

    q = lambda di, dj, dk: phi[i0[2] + dk, i0[1] + dj, i0[0] + di]
    e00, e10 = lerp(q(0, 0, 0), q(1, 0, 0), t[0]), lerp(q(0, 1, 0), q(1, 1, 0), t[0])
    e01, e11 = lerp(q(0, 0, 1), q(1, 0, 1), t[0]), lerp(q(0, 1, 1), q(1, 1, 1), t[0])
    return lerp(lerp(e00, e10, t[1]), lerp(e01, e11, t[1]), t[2])


def seeds(grid_size):
    phi, o, vox, d = snowball()
    h = f32(grid_size) / f32(2.0)
    lo = o
    hi = np.array([o[a] + vox * f32(d[a] - 1) for a in range(3)], dtype=f32)
    n = [int(np.ceil(f32(hi[a] - lo[a]) / h)) for a in range(3)]
    out = []
    for i in range(n[0]):
        for j in range(n[1]):
            for k in range(n[2]):
                p = np.array([lo[0] + f32(f32(i) + f32(0.5)) * h, lo[1] + f32(f32(j) + f32(0.5)) * h,
                              lo[2] + f32(f32(k) + f32(0.5)) * h], dtype=f32)
                if sample(phi, o, vox, d, p) < 0:
                    out.append(p)
    return np.array(out, dtype=f32)


def read_positions(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    return np.frombuffer(raw[4:], dtype=f32).reshape(n, 3)


def cook_binary():
    recipe = load_recipe()
    if not os.path.exists(recipe.COOK_BIN):
        recipe.build()
    return recipe.COOK_BIN


def test_seeding_matches_the_closed_form(tmp_path):
    out = str(tmp_path / "seed.bin")
    r = subprocess.run([cook_binary(), "seed", out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_positions(out)
    want = seeds(1.0 / 32.0)
    assert got.shape == want.shape and len(got) > 1000
    assert np.array_equal(got, want)
    # the seeds fill the ball at spacing gridSize / 2: about (4/3 pi r^3) / h^3 of them
    assert abs(len(got) - 4.0 / 3.0 * np.pi * 0.1 ** 3 * 64 ** 3) < 0.05 * len(got)
