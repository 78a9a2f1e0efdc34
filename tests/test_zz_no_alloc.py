"""The allocation-free entry point against the one it is a variant of, through the C-ABI.

mpm_model_stress_differential_noalloc takes device-resident particle indices, dF and output and
keeps its slot list in a fixed per-GPU device array. Its output must be bitwise that of
mpm_model_stress_differential (host arrays, buffers allocated per call) on the same particles: two
calls in a row, the second on fewer particles (the scratch keeps the first call's slots beyond
them), then the error paths: more particles than the scratch holds, and an index out of range.
Device memory for the caller's arrays comes from the HIP runtime directly (the caller allocates;
the entry point does not).
"""
import ctypes

import numpy as np
import pytest

import mpmsnow_amd as G

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    for name in ("libamdhip64.so.7", "libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = ctypes.CDLL(name)
        except OSError:
            continue
        hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        hip.hipFree.argtypes = [ctypes.c_void_p]
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        return hip
    raise RuntimeError("HIP runtime library not found")


class DevArray:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), ctypes.c_size_t(max(nbytes, 4))) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), H2D) == 0

    def get(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, ctypes.c_size_t(out.nbytes), D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


def _context(n, seed):
    rng = np.random.default_rng(seed)
    ctx = G.Context()
    ctx.particles_create(n)
    ctx.set("p", rng.uniform(0, 1, (n, 3)).astype(np.float32))
    ctx.set("m", np.full(n, 1e-3, np.float32))
    F = np.eye(3)[None] + 0.08 * rng.uniform(-1, 1, (n, 3, 3))
    ctx.set("F", F.astype(np.float32))
    ctx.model_snow(1.4e5, 0.2, 10.0, 2.5e-2, 7.5e-3)
    ctx.update_particle_data()
    return ctx, rng


def _noalloc(ctx, hip, idx, dF_cm):
    n = idx.size
    d_idx, d_dF, d_out = DevArray(hip, 4 * n), DevArray(hip, 36 * n), DevArray(hip, 36 * n)
    try:
        d_idx.put(idx.astype(np.int32))
        d_dF.put(dF_cm.astype(np.float32))
        d_out.put(np.zeros(9 * n, np.float32))
        rc = G.lib().mpm_model_stress_differential_noalloc(ctx.ptr, d_idx.p, d_dF.p, n, d_out.p)
        return rc, d_out.get(np.float32, 9 * n)
    finally:
        for d in (d_idx, d_dF, d_out):
            d.free()


def _alloc(ctx, idx, dF_cm):
    out = np.zeros(9 * idx.size, np.float32)
    ii = np.ascontiguousarray(idx, dtype=np.int32)
    d = np.ascontiguousarray(dF_cm, dtype=np.float32)
    rc = G.lib().mpm_model_stress_differential(ctx.ptr, ii.ctypes.data_as(G._ip), d.ctypes.data_as(G._fp),
                                               idx.size, out.ctypes.data_as(G._fp))
    assert rc == G.MPM_OK
    return out


def test_noalloc_stress_differential_matches_the_allocating_entry_point():
    print("[gpu-case] no-alloc stress differential: 2 calls (60000, then 777 particles) + error paths", flush=True)
    hip = _hip()
    n = 80000
    ctx, rng = _context(n, 3)
    for count in (60000, 777):  # the second call on fewer particles, in another order
        idx = rng.permutation(n)[:count].astype(np.int32)
        dF = rng.uniform(-1e-3, 1e-3, 9 * count).astype(np.float32)
        rc, got = _noalloc(ctx, hip, idx, dF)
        assert rc == G.MPM_OK, G.lib().mpm_last_error(ctx.ptr)
        want = _alloc(ctx, idx, dF)
        assert np.isfinite(want).all() and np.abs(want).max() > 0
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # more particles than the scratch holds: rejected on the host before any launch (the small
    # buffers passed are never read)
    small = DevArray(hip, 64)
    try:
        rc = G.lib().mpm_model_stress_differential_noalloc(ctx.ptr, small.p, small.p, (1 << 20) + 1, small.p)
    finally:
        small.free()
    assert rc == G.MPM_ERR_INVALID
    assert b"at most" in G.lib().mpm_last_error(ctx.ptr)
    # an index out of range: the call fails like the allocating one, and the next call works again
    idx = np.array([5, n, 7], np.int32)
    rc, got = _noalloc(ctx, hip, idx, np.zeros(27, np.float32))
    assert rc == G.MPM_ERR_INVALID
    assert b"out of range" in G.lib().mpm_last_error(ctx.ptr)
    assert np.isnan(got[9:18]).all()
    idx = np.array([5, 9, 7], np.int32)
    dF = rng.uniform(-1e-3, 1e-3, 27).astype(np.float32)
    rc, got = _noalloc(ctx, hip, idx, dF)
    assert rc == G.MPM_OK
    assert np.array_equal(got.view(np.uint32), _alloc(ctx, idx, dF).view(np.uint32))
