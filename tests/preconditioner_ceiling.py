"""How far any preconditioner of the learned preconditioner's family can cut the slab's CR iterations.

The learned preconditioner attaches where ConjugateResiduals takes a preconditioner
(ConjugateResiduals.h:20, ProceduralMatrix.h:16) and is measured against the reference's diagonal
one (Grid.cpp:836-886). This script bounds what that can gain on the slope-slab system, on the CPU,
with the FP64 oracle as the operator: the implicit matrix of Grid.cpp:727-832 of the largest body,
after --warmup steps of the slab cut to --n particles at the 4M slab's density, h, dt and materials,
solved to the reference's 1e-5 relative residual from the reference's initial guess (Grid.cpp:911-947)
by the reference's preconditioned CR recurrences with:

  * the diagonal preconditioner (the reference's);
  * 3x3 node-block Jacobi (each node's 3x3 block of A, probed with HVPs);
  * exact two-level preconditioners: piecewise-constant aggregation of 2x2x2 nodes, Galerkin coarse
    operator P^T A P (probed with HVPs) factorised exactly; additive z = D^-1 r + P Ac^-1 P^T r and
    multiplicative z = c + D^-1 (r - A c), c = P Ac^-1 P^T r (one more HVP per application);
  * the diagonal CR from a corrected initial guess: the exact coarse correction of r0, and
    x0 + f (x* - x0) for f = 0.5, 0.9, 0.99 (a corrector that removes the fraction f of the
    initial error in every mode).

    python tests/preconditioner_ceiling.py --n 30000 --warmup 6 --out <summary .md>

Test infrastructure: it uses the oracle as the operator and is not part of the library.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (ROOT, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)


def slab_system(n, warmup):
    """The implicit system of the largest body after `warmup` steps: dict with the operator A, the
    projection, rhs b, diagonal, initial guess x0 and the grid."""
    from oracle import pyoracle as O
    import precision_study as PS

    sc, _ = PS.study_scene(n)
    sim = PS._oracle(O, sc, "double")
    for _ in range(warmup):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    b0 = max(sim.bodies(), key=len)
    pd = sim.pd
    v, m = pd.get("v"), pd.get("m")
    cmv = (v[b0] * m[b0, None]).sum(0) / m[b0].sum()
    h, dt = float(np.float32(sc["h"])), sc["dt"]
    cols, fields, model = sim._keep
    g = O.Grid(pd, b0, h, frame_v=cmv)
    _, flags = g.explicit_momenta(dt, model, cols, fields)
    res = g.update_grid_velocities(dt, model, cols, fields, sc["max_iters"], sc["tol"], want_rhs=True)
    mass = g.get("masses")
    N = g.num_nodes
    diag = -dt * dt * g.dforce_dxi(model) + np.repeat(mass, 3)
    diag[diag == 0] = 1.0
    nrm = np.zeros((N, 3))
    nrm[flags >= 0] = sc["planes"][0][0][:3]  # the slab's one collider: the ground plane

    def project(x):
        x = x.reshape(N, 3).copy()
        x[flags == -2] = 0.0
        sel = flags >= 0
        x[sel] -= (x[sel] * nrm[sel]).sum(1)[:, None] * nrm[sel]
        return x.reshape(-1)

    def A(x):
        y = project(x)
        return project(np.repeat(mass, 3) * y - dt * dt * g.force_differentials(model, y))

    return dict(A=A, project=project, b=res["rhs"], diag=diag, x0=np.tile(res["guess"], N), grid=g, mass=mass,
                flags=flags, N=N, ref_iterations=res["iterations"], particles=len(b0))


def cr(S, apply_minv, x0, tol=1e-5, maxit=1000):
    """The reference's preconditioned CR (ConjugateResiduals.cpp:19-160) with the termination
    |b - A x|^2 < tol^2 |b|^2; returns (x updates, solution)."""
    A, project, b = S["A"], S["project"], S["b"]
    x = x0.copy()
    thr = tol * tol * (b @ b)
    r = b - A(x)
    rp = apply_minv(r)
    p = rp.copy()
    Ap = A(p)
    rz = rp @ Ap
    for i in range(maxit):
        pAp = apply_minv(Ap)
        a = rz / (Ap @ pAp)
        x = project(x + a * p)
        rp = rp - a * pAp
        r = b - A(x)
        if r @ r < thr:
            return i + 1, x
        Ar = A(rp)
        rzo, rz = rz, rp @ Ar
        p = rp + (rz / rzo) * p
        Ap = Ar + (rz / rzo) * Ap
    return maxit, x


def node_coords(S):
    n0, n1, _ = S["grid"].n
    idx = np.arange(S["N"])
    return idx % n0, (idx // n0) % n1, idx // (n0 * n1)


def block_jacobi(S):
    """Inverse 3x3 node blocks of A by probing (nodes 7 apart on every axis do not couple)."""
    N, A = S["N"], S["A"]
    i, j, k = node_coords(S)
    col = (i % 7) + 7 * (j % 7) + 49 * (k % 7)
    B = np.zeros((N, 3, 3))
    for c in np.unique(col):
        sel = col == c
        for a in range(3):
            e = np.zeros((N, 3))
            e[sel, a] = 1.0
            B[sel, :, a] = A(e.reshape(-1)).reshape(N, 3)[sel]
    Binv = np.empty_like(B)
    for q in range(N):
        if S["flags"][q] != -1 or not np.abs(B[q]).max() > 0:
            Binv[q] = np.diag(1.0 / S["diag"][3 * q:3 * q + 3])
        else:
            Binv[q] = np.linalg.inv(B[q])
    return lambda r: np.einsum("nij,nj->ni", Binv, r.reshape(N, 3)).reshape(-1)


def coarse_space(S, agg=2):
    """Aggregation P (2x2x2 nodes with mass -> one coarse node) and the factorised Galerkin operator."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    N, A = S["N"], S["A"]
    n0, n1, n2 = S["grid"].n
    i, j, k = node_coords(S)
    c0, c1 = (n0 + agg - 1) // agg, (n1 + agg - 1) // agg
    aid = (i // agg) + c0 * ((j // agg) + c1 * (k // agg))
    act = S["mass"] > 0
    used = np.unique(aid[act])
    remap = -np.ones(aid.max() + 1, dtype=np.int64)
    remap[used] = np.arange(len(used))
    Nc = len(used)
    idx = np.arange(N)[act]
    rows = np.concatenate([3 * idx + a for a in range(3)])
    cols = np.concatenate([3 * remap[aid[act]] + a for a in range(3)])
    P = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(3 * N, 3 * Nc))
    ai, aj, ak = used % c0, (used // c0) % c1, used // (c0 * c1)
    stride = 3 + (3 + agg - 1) // agg  # aggregates this far apart do not couple through A
    colour = (ai % stride) + stride * ((aj % stride) + stride * (ak % stride))
    Ac = sp.lil_matrix((3 * Nc, 3 * Nc))
    for c in np.unique(colour):
        sel = np.where(colour == c)[0]
        for a in range(3):
            e = np.zeros(3 * Nc)
            e[3 * sel + a] = 1.0
            y = P.T @ A(P @ e)
            for q in np.nonzero(y)[0]:
                J = q // 3
                d = np.maximum.reduce([np.abs(ai[sel] - ai[J]), np.abs(aj[sel] - aj[J]), np.abs(ak[sel] - ak[J])])
                Ac[q, 3 * sel[np.argmin(d)] + a] = y[q]
    Ac = sp.csc_matrix(Ac)
    Ac = 0.5 * (Ac + Ac.T)
    zero = np.where(np.abs(Ac).sum(1).A1 == 0)[0]  # aggregates of projected (collided) nodes only
    Ac = Ac + sp.csc_matrix((np.ones(len(zero)), (zero, zero)), shape=Ac.shape)
    lu = spla.splu(Ac)
    return P, lu, Nc


def run(n, warmup, log=print):
    t0 = time.time()
    S = slab_system(n, warmup)
    dinv = 1.0 / S["diag"]
    out = {"particles": S["particles"], "nodes": int(S["N"]), "active_nodes": int((S["mass"] > 0).sum()),
           "reference_iterations": S["ref_iterations"]}
    it_d, xs = cr(S, lambda r: r * dinv, S["x0"])
    out["diagonal"] = it_d
    log("diagonal CR: %d iterations (the oracle's own solve: %d)" % (it_d, S["ref_iterations"]))
    out["block_jacobi"] = cr(S, block_jacobi(S), S["x0"])[0]
    log("3x3 block Jacobi: %d" % out["block_jacobi"])
    P, lu, Nc = coarse_space(S)
    out["coarse_unknowns"] = 3 * Nc
    coarse = lambda r: P @ lu.solve(P.T @ r)  # noqa: E731
    out["two_level_additive"] = cr(S, lambda r: r * dinv + coarse(r), S["x0"])[0]

    def mult(r):
        c = coarse(r)
        return c + dinv * (r - S["A"](c))

    out["two_level_multiplicative"] = cr(S, mult, S["x0"])[0]
    log("two-level additive %d, multiplicative %d (coarse unknowns %d)" % (
        out["two_level_additive"], out["two_level_multiplicative"], 3 * Nc))
    b = S["b"]
    r0 = b - S["A"](S["x0"])
    x1 = S["project"](S["x0"] + coarse(r0))
    out["r0_over_b"] = float(np.linalg.norm(r0) / np.linalg.norm(b))
    out["coarse_guess_r_over_b"] = float(np.linalg.norm(b - S["A"](x1)) / np.linalg.norm(b))
    out["coarse_corrected_guess"] = cr(S, lambda r: r * dinv, x1)[0]
    out["exact_fraction_guess"] = {}
    for f in (0.5, 0.9, 0.99):
        out["exact_fraction_guess"][str(f)] = cr(S, lambda r: r * dinv, S["x0"] + f * (xs - S["x0"]))[0]
    log("coarse-corrected initial guess %d; x0 + f (x* - x0): %s" % (
        out["coarse_corrected_guess"], out["exact_fraction_guess"]))
    out["seconds"] = time.time() - t0
    return out


def summarize(res, n, warmup, out_path):
    f = res["exact_fraction_guess"]
    lines = [
        "# Preconditioner ceiling on the slope-slab system (CPU, FP64 oracle operator)",
        "",
        "Command: `python tests/preconditioner_ceiling.py --n %d --warmup %d --out <this file>`" % (n, warmup),
        "",
        "The implicit system of the slab's largest body (%d particles, %d grid nodes, %d with mass) after %d steps"
        % (res["particles"], res["nodes"], res["active_nodes"], warmup),
        "of the slab cut at the 4M slab's density, h = 1/256, dt = 5e-4 and materials; the reference's CR",
        "recurrences to a 1e-5 relative residual from the reference's initial guess (|r0| / |b| = %.2f)."
        % res["r0_over_b"],
        "",
        "| preconditioner / start | CR iterations | cost per iteration beyond the diagonal CR's |",
        "|---|---|---|",
        "| diagonal (the reference's) | %d | - |" % res["diagonal"],
        "| 3x3 node-block Jacobi | %d | a 3x3 product per node |" % res["block_jacobi"],
        "| exact two-level, additive (2x2x2 aggregates, Galerkin coarse operator factorised) | %d | a coarse solve |"
        % res["two_level_additive"],
        "| exact two-level, multiplicative (coarse solve, then one Jacobi step) | %d | a coarse solve and one HVP |"
        % res["two_level_multiplicative"],
        "| diagonal, from the exact coarse-corrected initial guess (|r| / |b| %.2f) | %d | one coarse solve per solve |"
        % (res["coarse_guess_r_over_b"], res["coarse_corrected_guess"]),
        "| diagonal, from x0 + 0.5 (x* - x0) | %d | - |" % f["0.5"],
        "| diagonal, from x0 + 0.9 (x* - x0) | %d | - |" % f["0.9"],
        "| diagonal, from x0 + 0.99 (x* - x0) | %d | - |" % f["0.99"],
        "",
        "%.0f s." % res["seconds"],
        "",
        "<!-- json: %s -->" % json.dumps(res),
    ]
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=30000)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    res = run(a.n, a.warmup, log=lambda s: print(s, flush=True))
    summarize(res, a.n, a.warmup, a.out)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
