"""Pins of the learned preconditioner's CPU oracle (oracle/lpc_oracle.py).

The network has no counterpart in the reference, so these checks come from the
mathematics: a brute-force loop convolution, adjoint identities of conv/pool,
and central finite differences of the float64 loss against the analytic
backward pass.
"""
import numpy as np
import pytest

from oracle import lpc_oracle as L


def brute_conv(x, W, b):
    X, Y, Z, _ = x.shape
    out = np.zeros((X, Y, Z, W.shape[0]))
    for i in range(X):
        for j in range(Y):
            for k in range(Z):
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            ii, jj, kk = i + dx, j + dy, k + dz
                            if 0 <= ii < X and 0 <= jj < Y and 0 <= kk < Z:
                                off = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)
                                out[i, j, k] += W[:, off, :] @ x[ii, jj, kk]
    return out + b


def test_conv_matches_brute_force_loops():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, 4, 2, 16))
    W = rng.normal(size=(16, 27, 16))
    b = rng.normal(size=16)
    assert np.allclose(L.conv(x, W, b), brute_conv(x, W, b), rtol=1e-12, atol=1e-12)


def test_conv_transpose_and_pool_are_adjoints():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(5, 3, 4, 16))
    W = rng.normal(size=(32, 27, 16))
    g = rng.normal(size=(5, 3, 4, 32))
    lhs = np.sum((L.conv(x, W, np.zeros(32))) * g)
    rhs = np.sum(x * L.conv_transpose(g, W))
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    y = rng.normal(size=(3, 2, 2, 16))
    assert abs(np.sum(L.pool_sum(x) * y) - np.sum(x * L.unpool(y, x.shape))) < 1e-10 * np.sum(np.abs(x))


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    r = L.round_bf16(vals)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0 and r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 2 ** -8 * 3.0e-3


def _problem(seed, n=(9, 7, 6)):
    rng = np.random.default_rng(seed)
    params = np.zeros(L.num_params())
    offs, _ = L.layer_offsets()
    for (cin, cout, _, _), (w, b) in zip(L.LAYERS, offs):
        params[w:w + 27 * cin * cout] = rng.normal(scale=np.sqrt(2.0 / (27 * cin)), size=27 * cin * cout)
        params[b:b + cout] = rng.normal(scale=0.05, size=cout)
    r = rng.normal(size=n + (3,))
    diag = rng.uniform(0.5, 2.0, size=n + (3,))
    mass = (rng.uniform(size=n) > 0.3).astype(np.float64)
    e = rng.normal(size=n + (3,))
    return params, r, diag, mass, e


def test_backward_matches_finite_differences():
    params, r, diag, mass, e = _problem(2)
    scale, weight = 1.3, 0.7
    loss, grad = L.loss_and_grad(params, r, diag, mass, e, scale, weight, bf16=False)
    rng = np.random.default_rng(3)
    offs, n = L.layer_offsets()
    idx = []
    for (cin, cout, _, mode), (w, b) in zip(L.LAYERS, offs):
        ncol = 3 if mode == "final" else cout  # only output channels 0..2 reach the loss
        for _ in range(4):
            co, off, ci = rng.integers(ncol), rng.integers(27), rng.integers(cin)
            idx.append(w + (co * 27 + off) * cin + ci)
        idx.append(b + rng.integers(ncol))
    eps = 1e-6
    for i in idx:
        p1 = params.copy()
        p1[i] += eps
        p2 = params.copy()
        p2[i] -= eps
        fd = (L.loss_and_grad(p1, r, diag, mass, e, scale, weight, bf16=False)[0] -
              L.loss_and_grad(p2, r, diag, mass, e, scale, weight, bf16=False)[0]) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_output_layer_channels_beyond_three_get_no_gradient():
    params, r, diag, mass, e = _problem(4, n=(5, 5, 5))
    _, grad = L.loss_and_grad(params, r, diag, mass, e, 1.0, 1.0, bf16=False)
    offs, _ = L.layer_offsets()
    w, b = offs[5]
    gW = grad[w:w + 27 * 16 * 16].reshape(16, 27, 16)
    assert np.all(gW[3:] == 0) and np.all(grad[b + 3:b + 16] == 0)
    assert np.abs(gW[:3]).max() > 0


def test_energy_loss_pins():
    """The exact-correction energy loss of oracle/lpc_oracle.py: 0 for any multiple of the exact
    correction, the error fraction a plain gradient step leaves (by direct evaluation of
    (e - alpha z)^T A (e - alpha z) at the minimising alpha), in [0, 1], and its gradient against
    central finite differences on a random SPD system."""
    rng = np.random.default_rng(4)
    n = 24
    M = rng.normal(size=(n, n))
    A = M @ M.T + n * np.eye(n)
    r = rng.normal(size=n)
    e = np.linalg.solve(A, r)
    E = float(e @ r)
    loss, _ = L.energy_loss_and_grad(-2.5 * e, r, A @ (-2.5 * e), E)
    assert abs(loss) < 1e-12
    z = rng.normal(size=n)
    loss, g = L.energy_loss_and_grad(z, r, A @ z, E)
    alphas = np.linspace(-5, 5, 20001) * (z @ r) / (z @ A @ z)
    left = min((e - al * z) @ A @ (e - al * z) for al in alphas)
    assert abs(loss - left / E) < 1e-6 and 0.0 <= loss <= 1.0
    h = 1e-6
    fd = np.array([(L.energy_loss_and_grad(z + h * np.eye(n)[i], r, A @ (z + h * np.eye(n)[i]), E)[0]
                    - L.energy_loss_and_grad(z - h * np.eye(n)[i], r, A @ (z - h * np.eye(n)[i]), E)[0]) / (2 * h)
                   for i in range(n)])
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.abs(g).max())
