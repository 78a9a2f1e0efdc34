"""Learned preconditioner on the GPU: network kernels against the CPU oracle
(oracle/lpc_oracle.py, itself pinned by finite differences in
test_lpc_oracle.py), training on pairs captured from the plain solver, and the
flexible-CR solve with fallback.

Tolerances: the GPU rounds activations and gradients to bfloat16 and sums in
FP32 (MFMA); the oracle rounds the forward activations at the same points and
sums in FP64, so outputs agree to a few bf16 ulps (2e-2 of the output scale)
and gradients to 5e-2 relative per layer.
"""
import numpy as np
import pytest

import mpmsnow_amd as G
from mpmsnow_amd import scenes
from oracle import lpc_oracle as L

pytestmark = pytest.mark.gpu


def say(msg):
    print("[lpc] " + msg, flush=True)


def grid_case(n=3000, seed=0):
    sc = scenes.snowball_drop(n=n, seed=seed)
    ctx = G.Context()
    ctx.particles_create(n)
    ctx.set("p", sc["x"])
    ctx.set("m", sc["m"])
    g = G.Grid(ctx, np.arange(n), sc["h"])
    return ctx, g, g.get("masses")


def random_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(L.num_params())
    offs, _ = L.layer_offsets()
    for (cin, cout, _, _), (w, b) in zip(L.LAYERS, offs):
        p[w:w + 27 * cin * cout] = rng.normal(scale=np.sqrt(2.0 / (27 * cin)), size=27 * cin * cout)
        p[b:b + cout] = rng.normal(scale=0.05, size=cout)
    return p.astype(np.float32)


def inputs(g, seed):
    rng = np.random.default_rng(seed)
    N = g.num_nodes
    r = rng.normal(size=(N, 3)).astype(np.float32)
    diag = rng.uniform(0.5, 2.0, size=(N, 3)).astype(np.float32)
    e = rng.normal(size=(N, 3)).astype(np.float32)
    return r, diag, e


def test_param_count_matches_oracle():
    assert G.lib().mpm_lpc_num_params() == L.num_params()


def test_network_forward_matches_oracle():
    say("forward snowball n=3000 seed=0")
    ctx, g, masses = grid_case()
    params = random_params(7)
    ctx.lpc_set_params(params)
    r, diag, _ = inputs(g, 8)
    scale = np.float32(1.7)
    out = g.lpc_apply(r, diag, scale)
    n = g.n
    ref, _ = L.forward(params, L.dense_to_volume(r, n, 3), L.dense_to_volume(diag, n, 3),
                       L.dense_to_volume(masses, n, 1), float(scale), bf16=True)
    ref = L.volume_to_dense(ref)
    err = np.abs(out - ref).max() / np.abs(ref).max()
    say("forward max err %.3e (scale %.3e)" % (err, np.abs(ref).max()))
    assert err < 2e-2
    assert np.abs(out).max() > 0


def test_network_gradients_match_oracle():
    say("gradients snowball n=2000 seed=1")
    ctx, g, masses = grid_case(n=2000, seed=1)
    params = random_params(11)
    ctx.lpc_set_params(params)
    r, diag, e = inputs(g, 12)
    scale, weight = np.float32(0.9), np.float32(1e-3)
    loss, grad = g.lpc_gradients(r, diag, e, scale, weight)
    n = g.n
    lc, gc = L.loss_and_grad(params, L.dense_to_volume(r, n, 3), L.dense_to_volume(diag, n, 3),
                             L.dense_to_volume(masses, n, 1), L.dense_to_volume(e, n, 3), float(scale),
                             float(weight), bf16=True)
    say("loss gpu %.6e cpu %.6e" % (loss, lc))
    assert abs(loss - lc) <= 1e-2 * abs(lc)
    offs, _ = L.layer_offsets()
    for k, ((cin, cout, _, _), (w, b)) in enumerate(zip(L.LAYERS, offs)):
        gw, cw = grad[w:w + 27 * cin * cout], gc[w:w + 27 * cin * cout]
        gb, cb = grad[b:b + cout], gc[b:b + cout]
        ew = np.linalg.norm(gw - cw) / np.linalg.norm(cw)
        eb = np.linalg.norm(gb - cb) / max(np.linalg.norm(cb), 1e-30)
        say("layer %d dW rel err %.3e db rel err %.3e" % (k, ew, eb))
        assert ew < 5e-2 and eb < 5e-2


def _small_scene(seed=3):
    sc = scenes.snowball_drop(n=20000, seed=seed, contact=True)
    sim = G.Sim(sc["x"], sc["m"], sc["h"], material=sc["material"], planes=sc["planes"], gravity=[sc["gravity"]],
                v=sc["v"])
    return sc, sim


def test_training_on_captured_pairs_reduces_loss():
    say("train: contact snowball 20k, capture 8 pairs")
    sc, sim = _small_scene()
    for _ in range(3):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    sim.ctx.lpc_init(5)
    sim.ctx.lpc_capture(8)
    sim.advance(sc["dt"], 400, 1e-7)
    first = sim.ctx.lpc_train(1, 1e-3)
    for _ in range(10):
        last = sim.ctx.lpc_train(1, 1e-3)
    say("loss %.4f -> %.4f" % (first, last))
    assert np.isfinite(last) and last < first


def test_learned_flexible_cr_meets_the_residual_test():
    """Untrained (zero output layer) network: every body falls back to the diagonal
    preconditioner after the first iteration and still meets |r|^2 < tol^2 |b|^2."""
    say("fcr: contact snowball 20k, untrained network")
    sc, sim = _small_scene(seed=4)
    for _ in range(2):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    sim.ctx.lpc_init(6)
    sim.ctx.lpc_mode(1)
    st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    assert st["num_bodies"] >= 1
    tol2 = np.float64(sc["tol"]) ** 2
    for b in range(st["num_bodies"]):
        s = sim.body_stats(b, hist_cap=0)
        if s["iterations"] == 0 or s["iterations"] > sc["max_iters"]:
            continue
        assert s["residual_norm2_last"] < tol2 * s["bnorm2"] or s["early_exit"], s
        assert s["fallbacks"] == 1
