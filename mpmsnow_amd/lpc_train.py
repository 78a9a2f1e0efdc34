"""Train the learned preconditioner on the GPU with the library's own kernels.

Runs seeded, randomised snow scenes (two colliding snowballs and slope slabs of
0.5-2M particles with varied radii, speeds, box sizes and Young's moduli) with
the diagonal-preconditioned CR solver; on chosen steps the solve keeps the
residuals r_k = b - A x_k of several CR iterations k and, before the particle
update, takes Adam steps of the network on them (mpm_lpc_train_next: forward,
residual loss |A z - r|^2_D through the matrix-free operator, backward on MFMA).
The exact correction A^-1 r_k is the loss's minimiser. The training seeds
(1000+) are disjoint from the evaluation scenes (seeds 1 and 2 of the benchmark
configurations).

    python -m mpmsnow_amd.lpc_train --out mpmsnow_amd/lpc_weights.npy

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
          decay=40.0, samples=12, capture_iters=None, offset=0):
    rng = np.random.default_rng(seed)
    params = None if init is None else np.load(init, allow_pickle=False)
    step = 0
    t0 = time.time()
    for k in range(scenes_n):
        name, sc, sim, warm = make_scene(k, rng, curriculum, offset)
        if capture_iters:
            sim.ctx.lpc_capture_iterations(capture_iters)
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
            step += 1
            print("[train] scene %d (%s) capture %d: loss %.4f -> %.4f (diagonal %.4f) lr %.2e (%.0f s)" % (
                k, name, c, first, last, diag, rate, time.time() - t0), flush=True)
            for _ in range(2):
                sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
        params = sim.ctx.lpc_params()
        np.save(out, params.astype(np.float32), allow_pickle=False)
        del sim
    return params


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
    a = ap.parse_args()
    its = [int(t) for t in a.capture_iters.split(",")] if a.capture_iters else None
    train(a.out, a.scenes, a.captures, a.epochs, a.lr, a.seed, a.init, a.curriculum, a.decay, a.samples, its,
          a.scene_offset)


if __name__ == "__main__":
    main()
