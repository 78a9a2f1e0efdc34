"""Seeded synthetic scenes shared by tests and bench.py.

The random stream is SplitMix64(seed) with u = (x >> 40) * 2**-24, so every
scene is reproducible bit-for-bit from its seed. Sphere samples use rejection
sampling in the bounding cube; boxes are uniform.
"""
import numpy as np

_M64 = (1 << 64) - 1


def splitmix64_uniform(seed, n):
    """n uniforms in [0,1) from SplitMix64(seed): u = (x >> 40) * 2^-24 (float32-exact)."""
    out = np.empty(n, dtype=np.float64)
    # vectorised SplitMix64: state_k = seed + (k+1)*golden
    k = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    out[:] = (z >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
    return out


def sphere(seed, n, center, radius):
    """n points uniform in a ball (rejection sampling), float32."""
    pts = []
    got = 0
    chunk = max(1024, int(n * 2.2))
    s = seed
    while got < n:
        u = splitmix64_uniform(s, 3 * chunk).reshape(chunk, 3)
        s += 1000003
        c = (2.0 * u - 1.0)
        keep = c[(c * c).sum(1) < 1.0]
        pts.append(keep)
        got += len(keep)
    p = np.concatenate(pts)[:n]
    return (np.asarray(center) + radius * p).astype(np.float32)


def box(seed, n, lo, hi):
    u = splitmix64_uniform(seed, 3 * n).reshape(n, 3)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    return (lo + (hi - lo) * u).astype(np.float32)


SNOW = dict(E=1.4e5, nu=0.2, xi=10.0, theta_c=2.5e-2, theta_s=7.5e-3)
RHO = 400.0


def snowball_drop(n=20000, seed=0, contact=False):
    """Snowball drop: n particles in a sphere r=0.1 at height 0.5, h=1/64, plane y=0 mu=0.2.

    contact=True is an added variant (centre y=0.2, v0=(0,-2,0)) that reaches the
    ground inside the run so the collision path is exercised.
    """
    r = 0.1
    c = (0.5, 0.2 if contact else 0.5, 0.5)
    x = sphere(seed, n, c, r)
    mp = RHO * (4.0 / 3.0 * np.pi * r ** 3) / n
    m = np.full(n, mp, dtype=np.float32)
    v = np.zeros((n, 3), dtype=np.float32)
    if contact:
        v[:, 1] = -2.0
    return dict(x=x, m=m, v=v, h=1.0 / 64, dt=1e-3, gravity=(0.0, -9.81, 0.0),
                planes=[((0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.2, False)], tol=1e-5, max_iters=1000,
                material=dict(SNOW))


def two_snowballs(n=1000000, seed=1, r=0.15, speed=3.0, E=None):
    """Two snowballs colliding: n/2 particles per sphere r=0.15 at (0.5 -/+ 0.16, 0.5, 0.5), v = (+/-3,0,0).

    r, speed and E vary the scene for training the learned preconditioner (the spheres stay
    0.02 m apart).
    """
    half = n // 2
    a = sphere(seed, half, (0.5 - (r + 0.01), 0.5, 0.5), r)
    b = sphere(seed + 7919, n - half, (0.5 + (r + 0.01), 0.5, 0.5), r)
    x = np.concatenate([a, b])
    mp = RHO * (4.0 / 3.0 * np.pi * r ** 3) / half
    m = np.full(n, mp, dtype=np.float32)
    v = np.zeros((n, 3), dtype=np.float32)
    v[:half, 0] = speed
    v[half:, 0] = -speed
    mat = dict(SNOW)
    if E is not None:
        mat["E"] = float(E)
    return dict(x=x, m=m, v=v, h=1.0 / 256, dt=5e-4, gravity=(0.0, -9.81, 0.0),
                planes=[((0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.2, False)], tol=1e-5, max_iters=1000,
                material=mat)


def slope_slab(n=4000000, seed=2, size=(0.8, 0.2, 0.4)):
    """Snow slab on a 30-degree slope, as tilted gravity over the plane y=0.

    Per-particle E0 log-uniform in [7e4, 2.8e5] Pa and xi uniform in [5, 10]. `size` varies the
    box for training the learned preconditioner.
    """
    sx, sy, sz = size
    x = box(seed, n, (0.5 - sx / 2, 0.0 + 1.0 / 256, 0.5 - sz / 2), (0.5 + sx / 2, sy, 0.5 + sz / 2))
    vol = sx * (sy - 1.0 / 256) * sz if size != (0.8, 0.2, 0.4) else 0.8 * 0.2 * 0.4
    m = np.full(n, RHO * vol / n, dtype=np.float32)
    u = splitmix64_uniform(seed + 17, 2 * n).reshape(n, 2)
    E0 = np.exp(np.log(7e4) + (np.log(2.8e5) - np.log(7e4)) * u[:, 0])
    xi = 5.0 + 5.0 * u[:, 1]
    nu = 0.2
    mu0 = (E0 / (2 * (1 + nu))).astype(np.float32)
    lam0 = (E0 * nu / ((1 + nu) * (1 - 2 * nu))).astype(np.float32)
    g = 9.81
    s30 = np.sin(np.pi / 6)
    c30 = np.cos(np.pi / 6)
    return dict(x=x, m=m, v=np.zeros((n, 3), dtype=np.float32), h=1.0 / 256, dt=5e-4,
                gravity=(g * s30, -g * c30, 0.0), planes=[((0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.2, False)],
                tol=1e-5, max_iters=1000, material=dict(SNOW), mu0=mu0, lambda0=lam0, xi=xi.astype(np.float32))


def stiff_drop(n=16000000, seed=4, h=1.0 / 512, v0=1.0):
    """Stiff stress case: 8 seeded boxes and spheres dropped onto the plane y=0 and an analytic sphere.

    Objects sit on a 4 x 2 layout (x centres 0.15 + i * 0.2333, z centres 0.33 / 0.67, +/- 5 mm
    jitter); each slot is a box (sides uniform in [0.10, 0.17] m) or a sphere (radius uniform in
    [0.06, 0.095] m), chosen by the seed. The sphere collider (centre (0.5, -0.05, 0.5), r = 0.25,
    friction 0.2) lies under the four inner slots; the outer four land on the ground plane
    (friction 0.2). Each object's lowest point starts 5-20 mm above the highest support point under
    its footprint and moves down at v0 (1 m/s), so every object is in contact within the first
    dozen steps of the 100. Particles per object are proportional to its volume (uniform density
    400 kg/m^3, about 6 particles per cell at h = 1/512). Snow with E0 = 4.8e5 Pa, xi = 10;
    h = 1/512, dt = 2e-3 (c dt / h = 35: stiff). Returns the scene dict plus `objects` (kind,
    centre, size, count per object).
    """
    u = splitmix64_uniform(seed, 64)
    sc_c, sc_r = np.array([0.5, -0.05, 0.5]), 0.25
    specs = []
    for s in range(8):
        i, kz = s % 4, s // 4
        cx = 0.15 + i * (0.7 / 3.0) + 0.01 * (u[8 * s] - 0.5)
        cz = (0.33 if kz == 0 else 0.67) + 0.01 * (u[8 * s + 1] - 0.5)
        if u[8 * s + 2] < 0.5:
            half = 0.5 * (0.10 + 0.07 * u[8 * s + 3:8 * s + 6])
            kind, vol, hx, hz = "box", float(np.prod(2 * half)), half[0], half[2]
        else:
            r = 0.06 + 0.035 * u[8 * s + 3]
            half = np.array([r, r, r])
            kind, vol, hx, hz = "sphere", 4.0 / 3.0 * np.pi * r ** 3, r, r
        # highest support under the footprint: the collider sphere at the footprint point nearest its axis
        qx = min(max(sc_c[0], cx - hx), cx + hx)
        qz = min(max(sc_c[2], cz - hz), cz + hz)
        d2 = (qx - sc_c[0]) ** 2 + (qz - sc_c[2]) ** 2
        support = max(0.0, sc_c[1] + np.sqrt(sc_r ** 2 - d2)) if d2 < sc_r ** 2 else 0.0
        cy = support + 0.005 + 0.015 * u[8 * s + 6] + half[1]
        specs.append((kind, (cx, cy, cz), half, vol))
    vols = np.array([sp[3] for sp in specs])
    counts = np.floor(n * vols / vols.sum()).astype(np.int64)
    counts[np.argmax(vols)] += n - counts.sum()
    xs, objects = [], []
    for s, ((kind, c, half, vol), k) in enumerate(zip(specs, counts)):
        if kind == "box":
            xs.append(box(seed * 100 + s, int(k), np.array(c) - half, np.array(c) + half))
        else:
            xs.append(sphere(seed * 100 + s, int(k), c, float(half[0])))
        objects.append(dict(kind=kind, center=tuple(float(t) for t in c), size=tuple(float(t) for t in 2 * half),
                            count=int(k)))
    x = np.concatenate(xs)
    m = np.concatenate([np.full(int(k), RHO * sp[3] / k, dtype=np.float32) for sp, k in zip(specs, counts)])
    v = np.zeros((n, 3), dtype=np.float32)
    v[:, 1] = -v0
    mat = dict(SNOW)
    mat["E"] = 4.8e5
    return dict(x=x, m=m, v=v, h=h, dt=2e-3, gravity=(0.0, -9.81, 0.0),
                planes=[((0.0, 1.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.2, False)],
                spheres=[(tuple(sc_c), sc_r, (0.0, 0.0, 0.0), 0.2, False)], tol=1e-5, max_iters=1000,
                material=mat, objects=objects)
