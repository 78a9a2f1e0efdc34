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


def grid_case(n=3000, seed=0, slab=False):
    # the snowball's grid has a few hundred level-0 tiles (the convolutions' four-wave form); the
    # slab's, 0.26 x 0.07 x 0.26 m at h = 1/256, about 1600 (the one-wave form of large levels)
    sc = (scenes.slope_slab(n=n, seed=seed, size=(0.26, 0.07, 0.26)) if slab
          else scenes.snowball_drop(n=n, seed=seed))
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


@pytest.mark.parametrize("case", ["snowball", "slab"])
def test_network_forward_matches_oracle(case):
    say("forward %s" % case)
    ctx, g, masses = grid_case() if case == "snowball" else grid_case(n=20000, seed=5, slab=True)
    say("level-0 tiles %d" % int(np.prod((np.asarray(g.n) + 3) // 4)))
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


@pytest.mark.parametrize("case", ["snowball", "slab"])
def test_network_gradients_match_oracle(case):
    say("gradients %s" % case)
    ctx, g, masses = grid_case(n=2000, seed=1) if case == "snowball" else grid_case(n=20000, seed=6, slab=True)
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
    """The contact snowball lowered to 5 mm above the ground, so that it is in contact from its
    third step on (in free fall the uniform guess solves the system at once)."""
    sc = scenes.snowball_drop(n=20000, seed=seed, contact=True)
    sc["x"] = (sc["x"] - np.float32([0.0, 0.095, 0.0])).astype(np.float32)
    sim = G.Sim(sc["x"], sc["m"], sc["h"], material=sc["material"], planes=sc["planes"], gravity=[sc["gravity"]],
                v=sc["v"])
    return sc, sim


def test_training_reduces_the_residual_loss_below_the_diagonal_step():
    """In-solve training on 8 residuals of one solve: the residual left by the preconditioned
    correction falls over the epochs and ends below that of the diagonal step alone."""
    say("train: contact snowball 20k, 8 residuals, 30 epochs")
    sc, sim = _small_scene()
    for _ in range(3):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    sim.ctx.lpc_init(5)
    sim.ctx.lpc_train_next(8, 30, 3e-4)
    sim.advance(sc["dt"], 400, 1e-7)
    first, last, diag = sim.ctx.lpc_last_loss()
    say("residual loss %.4f -> %.4f (diagonal step %.4f)" % (first, last, diag))
    assert np.isfinite(last) and last < first and last < diag


def test_exact_correction_training_reduces_the_energy_loss_below_the_diagonal_step():
    """Training on exact corrections (mpm_lpc_set_training_loss(1)): a tight solve keeps 8 pairs
    (r_k, x* - x_k); the error energy left by the best step along the network's correction falls
    over the epochs and ends below the diagonal step's. The loss setting is validated."""
    say("train exact-correction loss: contact snowball 20k, 8 pairs, 30 epochs")
    sc, sim = _small_scene()
    for bad in (-1, 2):
        with pytest.raises(G.MpmError):
            sim.ctx.lpc_training_loss(bad)
    for _ in range(3):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    sim.ctx.lpc_init(5)
    sim.ctx.lpc_training_loss(1)
    sim.ctx.lpc_train_next(8, 30, 3e-4)
    sim.advance(sc["dt"], 400, 1e-7)
    first, last, diag = sim.ctx.lpc_last_loss()
    say("energy loss %.4f -> %.4f (diagonal step %.4f)" % (first, last, diag))
    assert 0.0 <= last <= 1.0 and 0.0 < diag <= 1.0
    assert np.isfinite(last) and last < first and last < diag


def test_capture_iterations_setting():
    """The capture schedule is validated (1-16 increasing non-negative iterations); a training solve
    with a late schedule keeps residuals only at iterations it reached and trains on them."""
    say("train: capture schedule, contact snowball 20k")
    sc, sim = _small_scene()
    for bad in ([], [3, 3], [-1, 2], list(range(17))):
        with pytest.raises(G.MpmError):
            sim.ctx.lpc_capture_iterations(bad)
    for _ in range(3):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    sim.ctx.lpc_capture_iterations([0, 5, 10, 20, 40, 80, 160])
    sim.ctx.lpc_init(5)
    sim.ctx.lpc_train_next(7, 4, 1e-4)
    sim.advance(sc["dt"], 400, 1e-7)
    first, last, diag = sim.ctx.lpc_last_loss()
    say("residual loss %.4f -> %.4f (diagonal step %.4f)" % (first, last, diag))
    assert np.isfinite(first) and np.isfinite(last) and 0.0 <= diag <= 1.0


def _largest(sim, st):
    rows = [sim.body_stats(b, hist_cap=0) for b in range(st["num_solved"])]
    return max(rows, key=lambda r: r["particles"]), rows


@pytest.mark.parametrize("dirs", [2, 8])
def test_untrained_network_is_diagonal_cr(dirs):
    """Zero output layer: z = r / diag, so the flexible iteration in the diag^-1 inner product
    is preconditioned CR whatever the number of directions it keeps; it meets |r|^2 < tol^2 |b|^2
    in about the plain solver's iterations."""
    say("fcr: contact snowball 20k, untrained network (%d directions) vs plain CR" % dirs)
    its = []
    for mode in (0, 1):
        sc, sim = _small_scene(seed=4)
        for bad in (0, 9):
            with pytest.raises(G.MpmError):
                sim.ctx.lpc_directions(bad)
        sim.ctx.lpc_directions(dirs)
        for _ in range(2):
            sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
        sim.ctx.lpc_init(6)
        sim.ctx.lpc_mode(mode)
        st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
        big, rows = _largest(sim, st)
        tol2 = np.float64(sc["tol"]) ** 2
        for s in rows:
            if 0 < s["iterations"] <= sc["max_iters"]:
                assert s["residual_norm2_last"] < tol2 * s["bnorm2"] or s["early_exit"], s
        its.append(big["iterations"])
    say("largest body iterations: plain %d, untrained learned %d" % tuple(its))
    assert abs(its[1] - its[0]) <= max(2, its[0] // 10)


def test_bad_network_falls_back_to_diagonal():
    """A network with a large random output layer: the residual stops falling, the body switches
    to the diagonal preconditioner (for good after three fallbacks) and still meets the
    residual test."""
    say("fcr: contact snowball 20k, random network")
    sc, sim = _small_scene(seed=5)
    for _ in range(2):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    p = random_params(13)
    offs, _ = L.layer_offsets()
    w, b = offs[5]
    p[w:w + 27 * 16 * 16] *= 50.0
    sim.ctx.lpc_set_params(p)
    sim.ctx.lpc_mode(1)
    st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    big, _ = _largest(sim, st)
    say("largest body: %d iterations, fallbacks %d" % (big["iterations"], big["fallbacks"]))
    assert 1 <= big["fallbacks"] <= 3
    assert big["residual_norm2_last"] < np.float64(sc["tol"]) ** 2 * big["bnorm2"]


def test_fallback_ratio_setting():
    """The fallback ratio is validated (0 < ratio <= 1); at ratio 1 (a fallback only when |r|^2
    rises) a bad network still ends on the diagonal within three fallbacks and meets the test."""
    say("fcr: fallback ratio 1.0, contact snowball 20k, random network")
    sc, sim = _small_scene(seed=5)
    for bad in (0.0, 1.5):
        with pytest.raises(G.MpmError):
            sim.ctx.lpc_fallback(bad)
    sim.ctx.lpc_fallback(1.0)
    for _ in range(2):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    p = random_params(13)
    offs, _ = L.layer_offsets()
    w, b = offs[5]
    p[w:w + 27 * 16 * 16] *= 50.0
    sim.ctx.lpc_set_params(p)
    sim.ctx.lpc_mode(1)
    st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    big, _ = _largest(sim, st)
    say("largest body: %d iterations, fallbacks %d" % (big["iterations"], big["fallbacks"]))
    assert 1 <= big["fallbacks"] <= 3
    assert big["residual_norm2_last"] < np.float64(sc["tol"]) ** 2 * big["bnorm2"]


def test_handoff_to_cr_after_the_third_fallback():
    """A network with a large random output layer sends the body to the diagonal for good
    within a few iterations (here after three fallbacks, mpm_lpc_set_strikes(3); the strike
    count is validated). With the handoff (the default) the rest of the solve runs the CR
    recurrences from the flexible CR's iterate; without it, GCR(2) on the diagonal. Both meet
    the residual test from identical states; the handoff's history continues the flexible CR's
    (one entry per iteration) and it takes no more iterations than GCR(2) plus 10% (the two are
    the same Krylov iteration in exact arithmetic), and plain CR's plus the iterations before
    the handoff."""
    say("fcr handoff: contact snowball 20k, random network, 3 strikes, handoff on / off / plain CR")
    out = {}
    for name, mode, ho in (("handoff", 1, 1), ("gcr", 1, 0), ("plain", 0, 1)):
        sc, sim = _small_scene(seed=5)
        for bad in (-1, 2):
            with pytest.raises(G.MpmError):
                sim.ctx.lpc_handoff(bad)
        for bad in (0, 17):
            with pytest.raises(G.MpmError):
                sim.ctx.lpc_strikes(bad)
        sim.ctx.lpc_strikes(3)
        sim.ctx.lpc_handoff(ho)
        for _ in range(2):
            sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
        p = random_params(13)
        offs, _ = L.layer_offsets()
        w, b = offs[5]
        p[w:w + 27 * 16 * 16] *= 50.0
        sim.ctx.lpc_set_params(p)
        sim.ctx.lpc_mode(mode)
        st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
        big, rows = _largest(sim, st)
        big = sim.body_stats([r["particles"] for r in rows].index(big["particles"]))
        tol2 = np.float64(sc["tol"]) ** 2
        assert big["residual_norm2_last"] < tol2 * big["bnorm2"], (name, big["iterations"])
        assert len(big["residual_norm2"]) == big["iterations"]
        out[name] = big
    say("largest body iterations: handoff %d (fallbacks %d), GCR(2) %d, plain CR %d" % (
        out["handoff"]["iterations"], out["handoff"]["fallbacks"], out["gcr"]["iterations"], out["plain"]["iterations"]))
    assert out["handoff"]["fallbacks"] == 3
    assert out["handoff"]["iterations"] <= 1.1 * out["gcr"]["iterations"] + 2
    assert out["handoff"]["iterations"] <= out["plain"]["iterations"] + 16


def test_ablation_trained_network_against_plain_cr():
    """What the trained network gains, on a held-out scene. Trained in-solve on exact corrections
    (pairs (r_k, x* - x_k) of tight solves) of a stiff contact snowball (seed 7, E = 1.4e6 Pa), it
    is evaluated on another (seed 8): on the held-out solve's residuals the best step along its
    correction leaves less of the error energy than the diagonal step does (the gain asserted).
    The learned flexible CR and plain CR then run the held-out scene's next steps from identical
    states; both meet the residual test and their iteration counts are reported (measured, not
    assumed: the network's per-step gain does not shorten the Krylov solve on these systems)."""
    def scene(seed):
        sc = scenes.snowball_drop(n=20000, seed=seed, contact=True)
        sc["x"] = (sc["x"] - np.float32([0.0, 0.095, 0.0])).astype(np.float32)
        sc["material"]["E"] = 1.4e6
        sim = G.Sim(sc["x"], sc["m"], sc["h"], material=sc["material"], planes=sc["planes"],
                    gravity=[sc["gravity"]], v=sc["v"])
        return sc, sim
    sc, sim = scene(7)
    sim.ctx.lpc_init(9)
    sim.ctx.lpc_training_loss(1)
    for _ in range(2):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    for _ in range(3):
        sim.ctx.lpc_train_next(8, 30, 3e-4)
        sim.advance(sc["dt"], 400, 1e-7)
    params = sim.ctx.lpc_params()
    its, held = {}, None
    for mode in (0, 1):
        sc, sim = scene(8)
        sim.ctx.lpc_set_params(params)
        sim.ctx.lpc_training_loss(1)
        for _ in range(2):
            sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
        sim.ctx.lpc_train_next(8, 1, 0.0)  # evaluation only: the learning rate is zero
        sim.advance(sc["dt"], 400, 1e-7)
        first, last, diag = sim.ctx.lpc_last_loss()
        assert np.allclose(sim.ctx.lpc_params(), params)
        held = (first, diag, sim.ctx.lpc_last_pairs())
        sim.ctx.lpc_mode(mode)
        counts = []
        for _ in range(3):
            st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
            big, _ = _largest(sim, st)
            assert big["residual_norm2_last"] < np.float64(sc["tol"]) ** 2 * big["bnorm2"] or big["early_exit"]
            counts.append(big["iterations"])
        its[mode] = counts
    say("ablation, held-out scene: energy left by the step %.4f (network) vs %.4f (diagonal) over %d pairs; "
        "largest-body CR iterations plain %s, learned %s" % (held[0], held[1], held[2], its[0], its[1]))
    assert held[2] > 0 and held[0] < held[1]
    assert all(c > 0 for c in its[0] + its[1])
