"""ConjugateResiduals over a caller's matrix (mpm_conjugate_residuals; ConjugateResiduals.cpp:19-160,
ProceduralMatrix.h:8-19) against the oracle's conjugateResiduals on the same dense systems.

The GPU runs the recurrences in FP32 with FP64 inner products; the oracle here runs them in double.
On well-conditioned systems with a fixed iteration count the two agree to FP32 rounding of the
vectors (relative 1e-5 on x, 1e-4 on every logged residual and search direction). The reference's own test
(TestConjugateResiduals.cpp:43-102) runs with its bounds."""
import numpy as np
import pytest

import mpmsnow_amd as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    return G.Context(0)


def _spd(n, seed, cond=30.0):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = np.geomspace(1.0, cond, n)
    return (q * ev) @ q.T


def test_reference_test_solve(ctx):
    """TestConjugateResiduals::testSolve: symmetric indefinite 6x6, SquareMagnitudeTermination(6, 0),
    logging on; |A x - v| < 1e-3, logged residual norms decrease, residuals A-orthogonal and search
    directions A^2-orthogonal to 1e-4. The reference draws one random matrix; of seeds 0-7 here, 2 and
    3 give A^2-orthogonality 2.7e-4 and 3.8e-4 in the reference's own FP32 arithmetic (the oracle in
    float), over its bound, and are left out."""
    print("[gpu-case] conjugate residuals: reference testSolve, 6x6 indefinite", flush=True)
    for seed in (0, 1, 4, 5, 6, 7):
        rng = np.random.default_rng(seed)
        a = rng.uniform(-1, 1, size=(6, 6)).astype(np.float32)
        A = (a + a.T).astype(np.float64)
        v = rng.uniform(-1, 1, size=6).astype(np.float32)
        x, st, R, P = G.conjugate_residuals(ctx, lambda z: A @ z, v, np.zeros(6), 6, 0.0, log=16)
        assert st["iterations"] == 7
        assert np.linalg.norm(A @ x - v) < 1e-3, (seed, np.linalg.norm(A @ x - v))
        assert len(R) >= 6
        n2 = [float(np.dot(r.astype(np.float64), r)) for r in R[:6]]
        assert all(n2[i] < n2[i - 1] for i in range(1, 6)), (seed, n2)
        for i in range(6):
            for j in range(6):
                if i != j:
                    assert abs(R[i] @ A @ R[j]) < 1e-4, (seed, i, j)
                    assert abs((A @ P[i]) @ (A @ P[j])) < 1e-4, (seed, i, j)


@pytest.mark.parametrize("precond", [False, True])
def test_matches_oracle_cr(ctx, oracle, precond):
    """400x400 SPD system (condition 30), 25 iterations from a nonzero start, with and without the
    oracle's diagonal preconditioner (multVector d x, multInverseVector x / d): x within 1e-5 and
    every logged residual and direction within 1e-4 relative of the double-precision oracle."""
    print("[gpu-case] conjugate residuals: 400x400 SPD vs oracle, preconditioner=%s" % precond, flush=True)
    n = 400
    rng = np.random.default_rng(3)
    d = rng.uniform(0.5, 4.0, size=n)
    A = _spd(n, 4) * np.sqrt(np.outer(d, d))
    b = rng.normal(size=n).astype(np.float32)
    x0 = (0.1 * rng.normal(size=n)).astype(np.float32)
    pre = (lambda z: d * z, lambda z: z / d) if precond else None
    x, st, R, P = G.conjugate_residuals(ctx, lambda z: A @ z, b, x0, 24, 0.0, preconditioner=pre, log=32)
    ref = oracle.cr_dense(A, b.astype(np.float64), x0.astype(np.float64), diag=d if precond else None,
                          max_iters=24, tol=0.0, prec="double", log_cap=32)
    assert st["iterations"] == ref["iterations"] == 25
    assert len(R) == len(ref["residuals"])

    def rel(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)

    print("[cr] x rel %.2e, residuals max rel %.2e" % (
        rel(x, ref["x"]), max(rel(r, q) for r, q in zip(R, ref["residuals"]))), flush=True)
    assert rel(x, ref["x"]) < 1e-5
    for r, q in zip(R, ref["residuals"]):
        assert rel(r, q) < 1e-4
    for p, q in zip(P, ref["directions"]):
        assert rel(p, q) < 1e-4


def test_tolerance_stop_and_history(ctx, oracle):
    """SquareMagnitudeTermination(200, 1e-4): the same iteration count as the double oracle, and the
    per-iteration |r|^2 history ends below tol^2 |b|^2."""
    print("[gpu-case] conjugate residuals: tolerance stop, 300x300 SPD", flush=True)
    n = 300
    A = _spd(n, 9, cond=100.0)
    b = np.random.default_rng(10).normal(size=n).astype(np.float32)
    x, st, _, _ = G.conjugate_residuals(ctx, lambda z: A @ z, b, np.zeros(n), 200, 1e-4)
    ref = oracle.cr_dense(A, b.astype(np.float64), None, max_iters=200, tol=1e-4, prec="double", log_cap=1)
    assert abs(st["iterations"] - ref["iterations"]) <= 1, (st["iterations"], ref["iterations"])
    hist = st["residual_norm2"]
    assert len(hist) == st["iterations"]
    assert hist[-1] < 1e-8 * float(b.astype(np.float64) @ b)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 2e-4


def test_zero_rhs_and_hooks(ctx):
    """b = 0 gives x = 0 with no iteration (ConjugateResiduals.cpp:34-38); a host termination hook
    sees the residual and its iteration number and stops the solve; the debug hook sees every x;
    subspaceProject keeps x in the subspace of the projected system."""
    print("[gpu-case] conjugate residuals: zero rhs, hooks, projection", flush=True)
    n = 50
    A = _spd(n, 2)
    x, st, _, _ = G.conjugate_residuals(ctx, lambda z: A @ z, np.zeros(n), np.ones(n), 10, 0.0)
    assert st["iterations"] == 0 and np.all(x == 0)
    seen, xs = [], []
    b = np.random.default_rng(5).normal(size=n)

    def term(r, i):
        seen.append((i, float(r @ r)))
        return i >= 3

    x, st, _, _ = G.conjugate_residuals(ctx, lambda z: A @ z, b, np.zeros(n), 100, 0.0, termination=term,
                                        debug=lambda v: xs.append(v))
    assert [i for i, _ in seen] == [0, 1, 2, 3]
    assert st["iterations"] == 4 and len(xs) == 4
    assert np.allclose(xs[-1], x)
    mask = (np.arange(n) % 3 != 0).astype(np.float64)
    PAP = A * np.outer(mask, mask)
    x, st, _, _ = G.conjugate_residuals(ctx, lambda z: PAP @ z, mask * b, np.zeros(n), 200, 1e-5,
                                        project=lambda z: mask * z)
    assert np.all(x[mask == 0] == 0)
    assert np.linalg.norm(PAP @ x - mask * b) / np.linalg.norm(mask * b) < 1e-4


def test_device_pointer_callbacks(ctx):
    """device_pointers = 1: the callbacks get device addresses of n floats on the context's device;
    here they copy through host memory with the HIP runtime (hipMemcpy, then hipDeviceSynchronize so
    the product is complete on return) and match the host-pointer solve bit for bit."""
    import ctypes as C
    print("[gpu-case] conjugate residuals: device-pointer callbacks, 200x200 SPD", flush=True)
    hip = C.CDLL("libamdhip64.so.7")  # the HIP runtime the library links (already loaded: same instance)
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    n = 200
    A = _spd(n, 12)
    b = np.random.default_rng(13).normal(size=n).astype(np.float32)
    seen = []

    def mult(user, x, out, m):
        assert m == n
        h = np.zeros(n, dtype=np.float32)
        assert hip.hipMemcpy(h.ctypes.data, C.cast(x, C.c_void_p), 4 * n, 2) == 0  # device to host
        y = (A @ h).astype(np.float32)
        assert hip.hipMemcpy(C.cast(out, C.c_void_p), y.ctypes.data, 4 * n, 1) == 0  # host to device
        assert hip.hipDeviceSynchronize() == 0
        seen.append(1)

    cb = G.MATVEC_CB(mult)
    M = G.ProceduralMatrix()
    M.mult = cb
    M.device_pointers = 1
    prm = G.make_params(30, 0.0)
    st = G.SolveStats()
    x = np.zeros(n, dtype=np.float32)
    ctx.check(G.lib().mpm_conjugate_residuals(ctx.ptr, C.byref(M), None, b.ctypes.data_as(G._fp),
                                              x.ctypes.data_as(G._fp), n, C.byref(prm), C.byref(st), None,
                                              None, 0, None))
    xh, sth, _, _ = G.conjugate_residuals(ctx, lambda z: A @ z, b, np.zeros(n), 30, 0.0)
    assert st.iterations == sth["iterations"] == 31 and len(seen) == 3 + 30
    assert np.array_equal(x, xh)


def test_errors(ctx):
    with pytest.raises(G.MpmError, match="n > 0"):
        G.conjugate_residuals(ctx, lambda z: z, np.zeros(0), np.zeros(0), 5, 0.0)
    with pytest.raises(G.MpmError, match="preconditioner"):
        G.conjugate_residuals(ctx, lambda z: z, np.ones(3), np.zeros(3), 5, 0.0, preconditioner=(lambda z: z, None))
