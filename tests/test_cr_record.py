"""CR iterate stream (the reference's `debug.dat` record, TestGrid.cpp:398-453 / viewer.cpp:48-81):
the writer's bytes against an independent struct-packed construction of the same record, and
the reader on both."""
import struct

import numpy as np

from mpmsnow_amd.cr_record import ImplicitUpdateRecord, read_record


def packed_record(h, gmin, n, masses, ev, iterates):
    b = struct.pack("<f", h) + struct.pack("<3f", *gmin) + struct.pack("<3i", *n)
    b += struct.pack("<%df" % len(masses), *masses) + struct.pack("<%df" % len(ev), *ev)
    for it in iterates:
        b += struct.pack("<%df" % len(it), *it)
    return b


def test_record_bytes_match_the_reference_layout(tmp_path):
    rng = np.random.default_rng(3)
    n = (3, 2, 4)
    nn = 24
    masses = rng.uniform(0, 1, nn).astype(np.float32)
    masses[[1, 5, 17]] = 0
    em = rng.normal(size=3 * nn).astype(np.float32)
    vc = rng.normal(size=3 * nn).astype(np.float32)
    xs = [rng.normal(size=3 * nn).astype(np.float32) for _ in range(3)]
    path = tmp_path / "debug.dat"
    with ImplicitUpdateRecord(str(path), 0.25, (0.5, -1.0, 2.0), n, masses, em, vc) as rec:
        for x in xs:
            rec(x)
    # explicit velocities: momenta / mass where mass != 0, then + vc (TestGrid.cpp:412-424)
    ev = np.zeros(3 * nn, np.float32)
    for i in range(nn):
        if masses[i] != 0:
            ev[3 * i:3 * i + 3] = em[3 * i:3 * i + 3] / masses[i]
    ev = ev + vc
    expect = packed_record(0.25, (0.5, -1.0, 2.0), n, masses, ev, [x + vc for x in xs])
    assert path.read_bytes() == expect
    r = read_record(str(path))
    assert r["h"] == 0.25 and list(r["n"]) == list(n) and np.allclose(r["min"], (0.5, -1.0, 2.0))
    assert np.array_equal(r["masses"], masses) and np.array_equal(r["explicit_velocities"], ev)
    assert len(r["iterates"]) == 3 and all(np.array_equal(a, x + vc) for a, x in zip(r["iterates"], xs))
