"""CPU oracle of the learned preconditioner's network (forward and backward).

TEST INFRASTRUCTURE ONLY: loaded by tests/ to check the GPU kernels of
mpmsnow_amd/csrc/lpc_kernels.hip and lpc_net.hpp; never part of the product path.

The network has no counterpart in the reference (it attaches where
ConjugateResiduals takes a preconditioner, ConjugateResiduals.h:20). This file
restates it on dense per-body volumes with numpy:

    in  (16 ch, linear in q = clamp(r/diag/s): q (3), q m/diag (3), q collided (3), zeros)
    e1 = relu(conv0(in))                  level 0, 16 ch
    e2 = relu(conv1(pool(e1)))            level 1, 32 ch
    e3 = relu(conv2(pool(e2)))            level 2, 64 ch
    d2 = relu(conv3(unpool(e3))) + e2     level 1
    d1 = relu(conv4(unpool(d2))) + e1     level 0
    out = conv5(d1)[..., :3]

3x3x3 convolutions with zero padding outside the body's grid; level l+1 has
ceil(n/2) voxels per axis; pool = mean of the 8 children (missing children
count as zero), unpool = nearest parent. Parameters are one flat vector, per
layer W[cout][27 cin] (offset index (dx+1) + 3 (dy+1) + 9 (dz+1), channel
fastest) then b[cout]. With bf16=True the forward pass rounds inputs, weights
and every stored activation to bfloat16 where the GPU does; with bf16=False it
is plain float64 (used for the finite-difference check of the backward pass).
"""
import numpy as np

CLAMP = 8.0  # inputs and training targets, in units of the scale

LAYERS = [(16, 16, 0, "relu"), (16, 32, 1, "relu"), (32, 64, 2, "relu"), (64, 32, 1, "relu"), (32, 16, 0, "relu"),
          (16, 16, 0, "final")]


def layer_offsets():
    offs = []
    o = 0
    for cin, cout, _, _ in LAYERS:
        w = o
        o += 27 * cin * cout
        b = o
        o += cout
        offs.append((w, b))
    return offs, o


def num_params():
    return layer_offsets()[1]


def round_bf16(x):
    """Round float values to bfloat16 (nearest-even) and return them as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, bf16):
    offs, n = layer_offsets()
    assert len(params) == n
    out = []
    for (cin, cout, _, _), (w, b) in zip(LAYERS, offs):
        W = np.asarray(params[w:w + 27 * cin * cout], dtype=np.float64).reshape(cout, 27, cin)
        if bf16:
            W = round_bf16(W)
        out.append((W, np.asarray(params[b:b + cout], dtype=np.float64)))
    return out


def _shift(x, ox, oy, oz):
    """y[i,j,k] = x[i+ox, j+oy, k+oz] with zeros outside."""
    X, Y, Z = x.shape[:3]
    y = np.zeros_like(x)
    xs = slice(max(0, -ox), min(X, X - ox)), slice(max(0, -oy), min(Y, Y - oy)), slice(max(0, -oz), min(Z, Z - oz))
    src = slice(xs[0].start + ox, xs[0].stop + ox), slice(xs[1].start + oy, xs[1].stop + oy), \
        slice(xs[2].start + oz, xs[2].stop + oz)
    y[xs] = x[src]
    return y


def _offsets():
    for off in range(27):
        yield off, off % 3 - 1, (off // 3) % 3 - 1, off // 9 - 1


def conv(x, W, b):
    """out[v, co] = sum_off sum_ci W[co, off, ci] x[v + off, ci] + b[co]."""
    out = np.zeros(x.shape[:3] + (W.shape[0],))
    for off, ox, oy, oz in _offsets():
        out += np.tensordot(_shift(x, ox, oy, oz), W[:, off, :], axes=([3], [1]))
    return out + b


def conv_transpose(g, W):
    """Data gradient of conv: dx[v, ci] = sum_off sum_co W[co, off, ci] g[v - off, co]."""
    out = np.zeros(g.shape[:3] + (W.shape[2],))
    for off, ox, oy, oz in _offsets():
        out += np.tensordot(_shift(g, -ox, -oy, -oz), W[:, off, :], axes=([3], [0]))
    return out


def conv_dw(x, g):
    """dW[co, off, ci] = sum_v g[v, co] x[v + off, ci]."""
    dW = np.zeros((g.shape[3], 27, x.shape[3]))
    for off, ox, oy, oz in _offsets():
        dW[:, off, :] = np.tensordot(g, _shift(x, ox, oy, oz), axes=([0, 1, 2], [0, 1, 2]))
    return dW


def coarse_shape(shape):
    return tuple((s + 1) // 2 for s in shape[:3])


def pool_sum(x):
    """Sum of the 8 children (children outside the fine grid count as zero)."""
    X, Y, Z, C = x.shape
    cx, cy, cz = coarse_shape(x.shape)
    p = np.zeros((2 * cx, 2 * cy, 2 * cz, C))
    p[:X, :Y, :Z] = x
    return p.reshape(cx, 2, cy, 2, cz, 2, C).sum(axis=(1, 3, 5))


def unpool(x, fine_shape):
    y = np.repeat(np.repeat(np.repeat(x, 2, axis=0), 2, axis=1), 2, axis=2)
    return y[:fine_shape[0], :fine_shape[1], :fine_shape[2]]


def features(r, diag, mass, scale, bf16, collided=None):
    """r, diag: (X, Y, Z, 3); mass: (X, Y, Z); collided: (X, Y, Z) bool or None."""
    f = np.zeros(r.shape[:3] + (16,))
    inv = 1.0 / scale if scale > 0 else 0.0
    d32 = np.asarray(diag, np.float32)
    q = (np.asarray(r, np.float32) / d32).astype(np.float32)
    q = np.clip((q * np.float32(inv)).astype(np.float32), -CLAMP, CLAMP)
    m32 = np.asarray(mass, np.float32)
    f[..., :3] = q
    f[..., 3:6] = (q * (m32[..., None] / d32).astype(np.float32)).astype(np.float32)
    if collided is not None:
        f[..., 6:9] = q * np.asarray(collided, np.float32)[..., None]
    return round_bf16(f) if bf16 else f


def forward(params, r, diag, mass, scale, bf16=True):
    """Returns (out (X, Y, Z, 3), cache)."""
    R = round_bf16 if bf16 else (lambda a: a)
    Ws = unpack(params, bf16)
    c = {}
    c["in"] = features(r, diag, mass, scale, bf16)
    s0 = c["in"].shape
    c["e1"] = R(np.maximum(conv(c["in"], *Ws[0]), 0))
    c["p1"] = R(pool_sum(c["e1"]) * 0.125)
    c["e2"] = R(np.maximum(conv(c["p1"], *Ws[1]), 0))
    c["p2"] = R(pool_sum(c["e2"]) * 0.125)
    c["e3"] = R(np.maximum(conv(c["p2"], *Ws[2]), 0))
    c["u2"] = unpool(c["e3"], c["e2"].shape)
    c["a2"] = R(np.maximum(conv(c["u2"], *Ws[3]), 0))
    c["d2"] = R(c["a2"] + c["e2"])
    c["u1"] = unpool(c["d2"], s0)
    c["a1"] = R(np.maximum(conv(c["u1"], *Ws[4]), 0))
    c["d1"] = R(c["a1"] + c["e1"])
    out = conv(c["d1"], *Ws[5])[..., :3]
    return out, c


def loss_and_grad(params, r, diag, mass, e, scale, weight, bf16=True):
    """Loss weight * sum |out - clamp((e - r / diag) / scale)|^2 and its gradient (flat, like params).

    The solver uses z = r / diag + scale * out, so the network's target is the part of the
    exact correction e that the diagonal step misses.
    """
    Ws = unpack(params, bf16)
    out, c = forward(params, r, diag, mass, scale, bf16)
    inv = np.float32(1.0 / scale)
    q = (np.asarray(r, np.float32) / np.asarray(diag, np.float32)).astype(np.float32)
    t = np.clip(((np.asarray(e, np.float32) - q) * inv).astype(np.float32), -CLAMP, CLAMP)
    d = out - t.astype(np.float64)
    loss = weight * float(np.sum(d * d))
    offs, n = layer_offsets()
    grad = np.zeros(n)

    def put(k, dW, g):
        w, b = offs[k]
        cin, cout = LAYERS[k][:2]
        grad[w:w + 27 * cin * cout] = dW.reshape(-1)
        grad[b:b + cout] = g.sum(axis=(0, 1, 2))

    g6 = np.zeros(out.shape[:3] + (16,))
    g6[..., :3] = 2.0 * weight * d
    put(5, conv_dw(c["d1"], g6), g6)
    gd1 = conv_transpose(g6, Ws[5][0])
    ga1 = gd1 * (c["a1"] > 0)
    put(4, conv_dw(c["u1"], ga1), ga1)
    gu1 = conv_transpose(ga1, Ws[4][0])
    gd2 = pool_sum(gu1)
    ga2 = gd2 * (c["a2"] > 0)
    put(3, conv_dw(c["u2"], ga2), ga2)
    gu2 = conv_transpose(ga2, Ws[3][0])
    ge3 = pool_sum(gu2) * (c["e3"] > 0)
    put(2, conv_dw(c["p2"], ge3), ge3)
    gp2 = conv_transpose(ge3, Ws[2][0])
    ge2 = (gd2 + unpool(gp2, c["e2"].shape) * 0.125) * (c["e2"] > 0)
    put(1, conv_dw(c["p1"], ge2), ge2)
    gp1 = conv_transpose(ge2, Ws[1][0])
    ge1 = (gd1 + unpool(gp1, c["e1"].shape) * 0.125) * (c["e1"] > 0)
    put(0, conv_dw(c["in"], ge1), ge1)
    return loss, grad


def dense_to_volume(a, n, ncomp):
    """Reference node order i + n0 (j + n1 k) -> [i, j, k(, c)]."""
    a = np.asarray(a).reshape(n[2], n[1], n[0], ncomp) if ncomp > 1 else np.asarray(a).reshape(n[2], n[1], n[0])
    return a.transpose(2, 1, 0, 3) if ncomp > 1 else a.transpose(2, 1, 0)


def volume_to_dense(v):
    """[i, j, k, c] -> reference node order, (n, c)."""
    return np.ascontiguousarray(v.transpose(2, 1, 0, 3)).reshape(-1, v.shape[3])


def energy_loss_and_grad(z, r, Az, E):
    """Exact-correction (energy-norm) training loss of one body and its gradient in z.

    For a residual r with exact correction e (A e = r) and error energy E = e^T A e = e^T r, the
    step alpha z with the best alpha leaves (e - alpha z)^T A (e - alpha z) = E - (z^T r)^2 / (z^T A z),
    i.e. the fraction L = 1 - (z^T r)^2 / ((z^T A z) E) of the error energy. With a = z^T r and
    c = z^T A z (A symmetric): dL/dz = (2 a / (c E)) ((a / c) A z - r). The library's training
    kernels (k_lpc_energy_dots, k_lpc_energy_grad) evaluate these from the HVP u = A z."""
    z, r, Az = (np.asarray(v, dtype=np.float64).ravel() for v in (z, r, Az))
    a = float(z @ r)
    c = float(z @ Az)
    L = 1.0 - a * a / (c * E)
    g = (2.0 * a / (c * E)) * ((a / c) * Az - r)
    return L, g
