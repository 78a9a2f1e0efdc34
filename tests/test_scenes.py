"""Seeded scene generators (CPU): the stiff stress case's layout invariants."""
import numpy as np

from mpmsnow_amd import scenes


def test_stiff_drop_layout():
    sc = scenes.stiff_drop(n=80000, seed=4, h=1.0 / 128)
    x, m, objs = sc["x"], sc["m"], sc["objects"]
    assert x.shape == (80000, 3) and m.shape == (80000,)
    assert sum(o["count"] for o in objs) == 80000 and len(objs) == 8
    # reproducible from the seed
    sc2 = scenes.stiff_drop(n=80000, seed=4, h=1.0 / 128)
    assert np.array_equal(sc2["x"], x) and np.array_equal(sc2["m"], m)
    # uniform density: total mass = 400 kg/m^3 x total object volume
    vol = sum(np.pi / 6 * o["size"][0] ** 3 if o["kind"] == "sphere" else np.prod(o["size"]) for o in objs)
    assert abs(m.astype(np.float64).sum() - 400.0 * vol) < 1e-4 * 400.0 * vol
    # objects do not overlap (gaps of at least 2 h between their particles' boxes) and start above
    # their supports: the ground plane and the analytic sphere collider
    counts = np.cumsum([0] + [o["count"] for o in objs])
    boxes = [(x[counts[s]:counts[s + 1]].min(0), x[counts[s]:counts[s + 1]].max(0)) for s in range(8)]
    for a in range(8):
        for b in range(a + 1, 8):
            gap = np.maximum(boxes[b][0] - boxes[a][1], boxes[a][0] - boxes[b][1]).max()
            assert gap > 2 * sc["h"], (a, b, gap)
    assert x[:, 1].min() > 0.004
    (c, r, _, _, _), = sc["spheres"]
    d = np.sqrt(((x - np.asarray(c)) ** 2).sum(1)) - r
    assert d.min() > 0.004
    # the four inner slots sit over the sphere collider, the outer four over the plane
    for s, o in enumerate(objs):
        dxz = np.hypot(o["center"][0] - c[0], o["center"][2] - c[2])
        assert (dxz < r) == (s % 4 in (1, 2)), (s, dxz)
    assert np.all(sc["v"][:, 1] == -1.0)
