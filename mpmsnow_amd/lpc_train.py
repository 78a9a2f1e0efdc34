"""Train the learned preconditioner on the GPU with the library's own kernels.

Runs seeded, randomised snow scenes (two colliding snowballs and slope slabs of
0.5-2M particles with varied radii, speeds, box sizes and Young's moduli) with
the diagonal-preconditioned CR solver; on chosen steps the solve keeps the
residuals r_k = b - A x_k of several CR iterations k and, before the particle
update, takes Adam steps of the network on them (mpm_lpc_train_next: forward,
residual loss |A z - r|^2_D through the matrix-free operator, backward on MFMA).
The exact correction A^-1 r_k is the loss's minimiser. With --loss exact the
solve also keeps its iterates x_k and the network is trained on the pairs
(r_k, x* - x_k), x* the tight solve's final iterate, minimising the error
energy left by the best step along its output. The training seeds
(1000+) are disjoint from the evaluation scenes (seeds 1 and 2 of the benchmark
configurations).

    python -m mpmsnow_amd.lpc_train --out mpmsnow_amd/lpc_weights.npy
    python -m mpmsnow_amd.lpc_train --evaluate mpmsnow_amd/lpc_weights.npy

Prints one progress line per training step; writes float32 parameters in the
layout of mpm_lpc_get_params.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import mpmsnow_amd as G  # noqa: E402
from mpmsnow_amd import scenes  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "lpc_weights.npy")


def make_scene(k, rng, curriculum="mixed", offset=0):
    """Scene k of the training curriculum: "mixed" alternates snowball pairs and slope slabs,
    "stiff" draws stiff drops (8 boxes and spheres on the plane and the sphere collider,
    E0 = 4.8e5 Pa, h = 1/512, dt = 2e-3) of 6-10M particles."""
    seed = 1000 + offset + k
    if curriculum == "stiff":
        n = int(rng.integers(6, 11)) * 1000000
        sc = scenes.stiff_drop(n=n, seed=seed)
        sim = G.Sim(sc["x"], sc["m"], sc["h"], material=sc["material"], planes=sc["planes"], spheres=sc["spheres"],
                    gravity=[sc["gravity"]], v=sc["v"])
        return "stiff drop n=%d" % n, sc, sim, int(rng.integers(8, 16))
    if k % 2 == 0:
        n = int(rng.integers(5, 11)) * 100000
        sc = scenes.two_snowballs(n=n, seed=seed, r=float(rng.uniform(0.1, 0.16)), speed=float(rng.uniform(1.5, 4.0)),
                                  E=float(np.exp(rng.uniform(np.log(7e4), np.log(2.8e5)))))
        sim = G.Sim(sc["x"], sc["m"], sc["h"], material=sc["material"], planes=sc["planes"], gravity=[sc["gravity"]],
                    v=sc["v"])
        return "snowballs n=%d" % n, sc, sim, int(rng.integers(8, 40))
    n = int(rng.integers(5, 21)) * 100000
    size = (float(rng.uniform(0.4, 0.8)), float(rng.uniform(0.1, 0.2)), float(rng.uniform(0.2, 0.4)))
    sc = scenes.slope_slab(n=n, seed=seed, size=size)
    sim = G.Sim(sc["x"], sc["m"], sc["h"], material=sc["material"], planes=sc["planes"], gravity=[sc["gravity"]],
                mu0=sc["mu0"], lambda0=sc["lambda0"], xi=sc["xi"])
    return "slab n=%d size=%s" % (n, tuple(round(v, 2) for v in size)), sc, sim, 2


def train(out=DEFAULT_OUT, scenes_n=24, captures=4, epochs=8, lr=2e-4, seed=0, init=None, curriculum="mixed",
          decay=40.0, samples=12, capture_iters=None, offset=0, loss="residual", time_limit=None):
    rng = np.random.default_rng(seed)
    params = None if init is None else np.load(init, allow_pickle=False)
    step = 0
    pairs = 0
    t0 = time.time()
    for k in range(scenes_n):
        name, sc, sim, warm = make_scene(k, rng, curriculum, offset)
        if capture_iters:
            sim.ctx.lpc_capture_iterations(capture_iters)
        sim.ctx.lpc_training_loss(1 if loss == "exact" else 0)
        if params is None:
            sim.ctx.lpc_init(seed + 1)
        else:
            sim.ctx.lpc_set_params(params)
        for _ in range(warm):
            sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
        for c in range(captures):
            rate = lr * (0.5 ** (step / decay))
            sim.ctx.lpc_train_next(samples, epochs, rate)
            sim.advance(sc["dt"], 300, 1e-7)
            first, last, diag = sim.ctx.lpc_last_loss()
            pairs += sim.ctx.lpc_last_pairs()
            step += 1
            print("[train] scene %d (%s) capture %d: loss %.4f -> %.4f (diagonal %.4f) lr %.2e, %d pairs (%.0f s)" % (
                k, name, c, first, last, diag, rate, pairs, time.time() - t0), flush=True)
            for _ in range(2):
                sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
        params = sim.ctx.lpc_params()
        np.save(out, params.astype(np.float32), allow_pickle=False)
        del sim
        if time_limit is not None and time.time() - t0 > time_limit:
            print("[train] time limit reached after scene %d: %d pairs" % (k, pairs), flush=True)
            break
    return params


def evaluate(weights, n=4000000, warmup=5, captures=3, samples=16):
    """Held-out per-step losses of a network: the 4M slope slab (seed 2, never a training scene)
    after `warmup` plain steps; each of `captures` steps keeps `samples` residuals of a tight solve
    and evaluates, without training (learning rate 0), the loss of the network's step and of the
    diagonal step, under both losses. Returns the means."""
    sc = scenes.slope_slab(n=n)
    sim = G.Sim(sc["x"], sc["m"], sc["h"], material=sc["material"], planes=sc["planes"], gravity=[sc["gravity"]],
                mu0=sc["mu0"], lambda0=sc["lambda0"], xi=sc["xi"])
    sim.ctx.lpc_set_params(np.load(weights, allow_pickle=False))
    for _ in range(warmup):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    out = {}
    for mode, name in ((0, "residual"), (1, "exact")):
        sim.ctx.lpc_training_loss(mode)
        net, diag, pairs = [], [], 0
        for _ in range(captures):
            sim.ctx.lpc_train_next(samples, 1, 0.0)
            sim.advance(sc["dt"], 300, 1e-7)
            first, _, d = sim.ctx.lpc_last_loss()
            net.append(first)
            diag.append(d)
            pairs += sim.ctx.lpc_last_pairs()
        out[name] = {"network": float(np.mean(net)), "diagonal": float(np.mean(diag)), "pairs": pairs}
        print("[evaluate] %s loss on the held-out 4M slab: network %.4f, diagonal step %.4f (%d pairs)" % (
            name, out[name]["network"], out[name]["diagonal"], pairs), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--scenes", type=int, default=24)
    ap.add_argument("--captures", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from these parameters (.npy)")
    ap.add_argument("--curriculum", choices=("mixed", "stiff"), default="mixed")
    ap.add_argument("--decay", type=float, default=40.0, help="learning-rate half-life in training steps")
    ap.add_argument("--samples", type=int, default=12, help="residuals kept per captured solve (at most 16)")
    ap.add_argument("--capture-iters", default=None,
                    help="comma-separated CR iterations whose residuals a training solve keeps")
    ap.add_argument("--scene-offset", type=int, default=0, help="scene k uses seed 1000 + offset + k")
    ap.add_argument("--loss", choices=("residual", "exact"), default="residual",
                    help="residual: the optimal step's residual |r - alpha A z|_D; exact: the error energy the "
                         "optimal step leaves of the exact correction x* - x_k (mpm_lpc_set_training_loss)")
    ap.add_argument("--time-limit", type=float, default=None, help="stop after the scene that passes this many seconds")
    ap.add_argument("--evaluate", default=None, metavar="WEIGHTS",
                    help="only evaluate these parameters on the held-out 4M slab under both losses")
    a = ap.parse_args()
    if a.evaluate:
        evaluate(a.evaluate)
        return
    its = [int(t) for t in a.capture_iters.split(",")] if a.capture_iters else None
    train(a.out, a.scenes, a.captures, a.epochs, a.lr, a.seed, a.init, a.curriculum, a.decay, a.samples, its,
          a.scene_offset, a.loss, a.time_limit)


if __name__ == "__main__":
    main()
