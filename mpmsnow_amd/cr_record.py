"""CR iterate stream: the reference's implicit-update record (`debug.dat`).

Format (little-endian, as TestGrid.cpp:398-453 writes it and viewer.cpp:48-81 reads it):

    float h; float min[3]; int n[3];
    float masses[n0 n1 n2];                 node order i + n0 (j + n1 k)
    float explicit_velocities[3 n0 n1 n2];  momenta / mass (0 where mass is 0) + collision velocities
    then, once per CR iteration, float iterate[3 n0 n1 n2] = x + collision velocities

The writer is a debug hook for `Grid.update_grid_velocities(..., debug=record)` (or
`Sim.advance`), which hands it each CR iterate x in the reference's node order. It is the
capture format for residual-history studies and preconditioner training pairs.
"""
import numpy as np


class ImplicitUpdateRecord:
    """Writes the header at construction and one iterate per call (TestGrid.cpp:398-453)."""

    def __init__(self, path, h, gmin, n, masses, explicit_momenta, collision_velocities):
        n = [int(v) for v in n]
        nn = n[0] * n[1] * n[2]
        masses = np.asarray(masses, dtype=np.float32).reshape(nn)
        em = np.asarray(explicit_momenta, dtype=np.float32).reshape(nn, 3)
        self.vc = np.asarray(collision_velocities, dtype=np.float32).reshape(3 * nn)
        ev = np.zeros((nn, 3), dtype=np.float32)
        nz = masses != 0
        ev[nz] = em[nz] / masses[nz, None]
        ev = ev.reshape(3 * nn) + self.vc
        self.f = open(path, "wb")
        self.f.write(np.float32(h).tobytes())
        self.f.write(np.asarray(gmin, dtype=np.float32).reshape(3).tobytes())
        self.f.write(np.asarray(n, dtype=np.int32).tobytes())
        self.f.write(masses.tobytes())
        self.f.write(ev.astype(np.float32).tobytes())
        self.count = 0

    @classmethod
    def for_grid(cls, path, grid, h, dt):
        """Header from a `Grid`: masses, explicit momenta at dt and the collision velocities of
        its collision flags (TestGrid.cpp:403-438)."""
        em, collided = grid.explicit_momenta(dt)
        vc = grid.collision_velocities(collided)
        return cls(path, h, grid.min, grid.n, grid.get("masses"), em, vc)

    def __call__(self, x):
        it = (np.asarray(x, dtype=np.float32) + self.vc).astype(np.float32)
        self.f.write(it.tobytes())
        self.count += 1

    def close(self):
        if self.f:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_record(path):
    """Reads a record (viewer.cpp:48-81): h, min, n, masses, explicit velocities, iterates."""
    with open(path, "rb") as fh:
        data = fh.read()
    o = 0

    def take(dtype, count):
        nonlocal o
        size = np.dtype(dtype).itemsize * count
        if o + size > len(data):
            raise ValueError("truncated record")
        a = np.frombuffer(data, dtype=dtype, count=count, offset=o)
        o += size
        return a

    h = float(take("<f4", 1)[0])
    gmin = take("<f4", 3).copy()
    n = take("<i4", 3).copy()
    nn = int(n[0]) * int(n[1]) * int(n[2])
    masses = take("<f4", nn).copy()
    ev = take("<f4", 3 * nn).copy()
    its = []
    step = 3 * nn * 4
    while o + step <= len(data):
        its.append(take("<f4", 3 * nn).copy())
    if o != len(data):
        raise ValueError("record ends inside an iterate")
    return {"h": h, "min": gmin, "n": n, "masses": masses, "explicit_velocities": ev, "iterates": its}
