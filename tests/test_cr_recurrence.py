"""The conjugate-residual denominator recurrence of the GPU kernels, checked on the CPU.

The reference computes the step length's denominator <Ap, M^-1 Ap> directly from the updated
Ap (ConjugateResiduals.cpp:86-102, :147-151). The GPU kernels take it from the pass that forms
Ar = A z by expanding <Ar + beta Ap, M^-1 (Ar + beta Ap)>, so that the p / Ap update shares a
pass with the next x update; the pass that stores the updated Ap also sums <Ap, M^-1 Ap> over the
stored vector, which the next expansion uses for its beta^2 term, and a non-positive expansion
restarts the recurrence (beta = 0). This file restates both forms with the GPU's precision
(FP32 vectors, FP64 inner products, alpha and beta rounded to FP32 where the vector updates use
them) and compares them on symmetric systems the diagonal-preconditioned CR finds hard:
positive definite of condition 1e2-1e8 and indefinite (which the MPM operator becomes where a
particle's F inverts). The restated recurrence without the resync is kept as the planted
mistake: it diverges from the direct form on indefinite systems, so the comparison can see the
difference the resync makes.
"""
import numpy as np
import pytest

f32 = np.float32
f64 = np.float64


def _system(n, cond, seed, negfrac):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0.0, np.log10(cond), n)
    ev[rng.random(n) < negfrac] *= -1.0
    a = (q * ev) @ q.T
    s = np.exp(rng.uniform(-1.0, 1.0, n))  # a non-trivial diagonal
    a = (a * s).T * s
    a = 0.5 * (a + a.T)
    b = np.random.default_rng(100 + seed).standard_normal(n).astype(f32)
    return a, np.abs(np.diag(a)).astype(f32), b


def _mv(a, v):
    return (a @ v.astype(f64)).astype(f32)


def _dot(u, v):
    return float(np.dot(u.astype(f64), v.astype(f64)))


def cr_direct(a, b, d, tol, maxit):
    """ConjugateResiduals.cpp:50-151 with x0 = 0, M = diag(d); returns |r|^2 per x update."""
    x = np.zeros_like(b)
    z = (b / d).astype(f32)
    p = z.copy()
    ap = _mv(a, p)
    rz = _dot(z, ap)
    thr = tol * tol * _dot(b, b)
    hist = []
    for i in range(maxit + 1):
        pap = (ap / d).astype(f32)
        den = _dot(ap, pap)
        if den == 0.0:
            break
        al = f32(rz / den)
        x = (x + al * p).astype(f32)
        z = (z - al * pap).astype(f32)
        r = (z * d).astype(f32)
        hist.append(_dot(r, r))
        if i >= maxit or hist[-1] < thr:
            break
        ar = _mv(a, z)
        rz_old, rz = rz, _dot(z, ar)
        if rz_old == 0.0:
            break
        be = f32(rz / rz_old)
        p = (z + be * p).astype(f32)
        ap = (ar + be * ap).astype(f32)
    return np.array(hist)


def cr_expanded(a, b, d, tol, maxit, resync=True, keep=1e-10):
    """The GPU kernels' sequence: k_cr_setup1/2, k_cr_update_x, then per iteration k_cr_ar
    (Ar, <z,Ar>, <Ar,M^-1 Ar>, <Ar,M^-1 Ap>; beta; expanded denominator; alpha) and
    k_cr_update_px (p, Ap, the stored Ap's <Ap,M^-1 Ap> when resync, x, z, |r|^2)."""
    x = np.zeros_like(b)
    z = (b / d).astype(f32)
    p = z.copy()
    ap = _mv(a, z)
    rz = _dot(z, ap)
    apdpap = _dot(ap, (ap / d).astype(f32))
    thr = tol * tol * _dot(b, b)
    al, be, hist, restarts = rz / apdpap, 0.0, [], 0
    for i in range(maxit + 1):
        if i > 0:
            p = (z + f32(be) * p).astype(f32)
            ap = (ar + f32(be) * ap).astype(f32)
            if resync:
                apdpap = _dot(ap, (ap / d).astype(f32))
        x = (x + f32(al) * p).astype(f32)
        z = (z - f32(al) * (ap / d).astype(f32)).astype(f32)
        r = (z * d).astype(f32)
        hist.append(_dot(r, r))
        if i >= maxit or hist[-1] < thr:
            break
        ar = _mv(a, z)
        t0, t1, t2 = _dot(z, ar), _dot(ar, (ar / d).astype(f32)), _dot(ar, (ap / d).astype(f32))
        rz_old, rz = rz, t0
        if rz_old == 0.0:
            break
        be = t0 / rz_old
        den = t1 + 2.0 * be * t2 + be * be * apdpap
        if not den > keep * (t1 + be * be * apdpap):
            be, den = 0.0, t1
            restarts += 1
        apdpap = den
        if den == 0.0:
            break
        al = rz / den
    return np.array(hist), restarts


CASES = [(cond, seed, 0.0) for cond in (1e2, 1e4, 1e6, 1e8) for seed in range(2)] + \
        [(cond, seed, 0.3) for cond in (1e1, 1e2, 1e4) for seed in range(4)]


def _spread(a, b, d, runs=4):
    """The direct form's own rounding spread: runs with every entry of b moved by one ulp."""
    rng = np.random.default_rng(7)
    out = []
    for _ in range(runs):
        bp = (b * (1.0 + rng.choice([-1.0, 1.0], b.size) * 2.0 ** -23)).astype(f32)
        out.append(cr_direct(a, bp, d, 1e-5, 1000))
    return out


@pytest.mark.parametrize("cond,seed,negfrac", CASES)
def test_expanded_denominator_tracks_the_direct_form(cond, seed, negfrac):
    a, d, b = _system(240, cond, seed, negfrac)
    bb = _dot(b, b)
    hd = cr_direct(a, b, d, 1e-5, 1000)
    he, restarts = cr_expanded(a, b, d, 1e-5, 1000)
    hp = _spread(a, b, d)
    assert restarts == 0
    k = min(10, len(hd), len(he), *[len(h) for h in hp])
    dev = np.max(np.abs(he[:k] - hd[:k]) / hd[:k])
    spread = max(np.max(np.abs(h[:k] - hd[:k]) / hd[:k]) for h in hp)
    fin = [h[-1] / bb for h in hp] + [hd[-1] / bb]
    its = [len(h) for h in hp] + [len(hd)]
    if negfrac == 0.0:
        # positive definite: the first iterations agree to 1e-5 (to rounding), the iteration count
        # and the final residual stay within the direct form's spread under one-ulp changes of b
        assert dev < 1e-5
        assert min(its) - max(8, 0.06 * min(its)) <= len(he) <= max(its) + max(8, 0.06 * max(its))
        assert min(fin) / 2.0 <= he[-1] / bb <= 2.0 * max(fin)
    else:
        # indefinite: rounding grows fast in both forms; the recurrence stays within 3x the direct
        # form's spread over the first iterations, converges when it converges, and otherwise
        # ends within 100x of its final residuals (the form without the resync: up to 1e4x)
        assert dev <= 3.0 * spread + 1e-6
        if max(fin) < 1e-10:
            assert he[-1] / bb < 1e-10
        else:
            assert he[-1] / bb <= 100.0 * max(fin)


def test_without_the_resync_the_recurrence_drifts_on_indefinite_systems():
    """The planted mistake: the previous denominator taken from the recurrence instead of the
    stored Ap. On these indefinite systems it ends orders of magnitude above the direct form's
    rounding spread, which the resynced recurrence stays within."""
    worst_old = worst_new = 0.0
    for cond in (1e2, 1e4):
        for seed in range(4):
            a, d, b = _system(240, cond, seed, 0.3)
            top = max(h[-1] for h in _spread(a, b, d) + [cr_direct(a, b, d, 1e-5, 1000)])
            ho, _ = cr_expanded(a, b, d, 1e-5, 1000, resync=False, keep=-np.inf)
            he, _ = cr_expanded(a, b, d, 1e-5, 1000)
            worst_old = max(worst_old, ho[-1] / top)
            worst_new = max(worst_new, he[-1] / top)
    assert worst_old > 300.0
    assert worst_new < 100.0
