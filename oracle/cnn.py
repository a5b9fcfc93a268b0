"""CPU oracle of the learned coarse-grid correction network (test infrastructure only).

The network has no counterpart in the reference.  It is a bias-free 3-level
U-Net on a periodic grid with channels-last tensors (nb, nz, ny, nx, C):

    L0  conv 3x3x3  cin -> c1, ReLU           a1
    L1  conv        c1  -> c1, ReLU           s1 (skip 1)
        avg-pool 2x2x2                        p1
    L2  conv        c1  -> c2, ReLU           a2
    L3  conv        c2  -> c2, ReLU           s2 (skip 2)
        avg-pool                              p2
    L4  conv        c2  -> c3, ReLU           a3
    L5  conv        c3  -> c3, ReLU           b3
        cat2 = [nearest-upsample(b3), s2]     (c3 + c2 channels)
    L6  conv        c3+c2 -> c2, ReLU         d2
    L7  conv        c2  -> c2, ReLU           u2
        cat1 = [upsample(u2), s1]             (c2 + c1 channels)
    L8  conv        c2+c1 -> c1, ReLU         d1
    L9  conv        c1  -> c1, ReLU           u1
    L10 conv        c1  -> cout (linear)      y

A convolution is the periodic cross-correlation
    Y[x, co] = sum_{tap, ci} W[co, tap, ci] * X[x + off(tap), ci],
tap = (dz+1)*9 + (dy+1)*3 + (dx+1), offsets in {-1, 0, 1}.  The parameters are
the weights of L0..L10 in that order, each W[co][tap][ci] flattened.  The loss
of a batch is sum((y - t)^2) / (2 * y.size).

With bf16=True the values are rounded to bfloat16 (round to nearest even) at
the points where the gfx950 kernels store them: the input, the weights, every
stored activation and every stored activation gradient; the arithmetic stays
FP64.  With bf16=False everything is FP64 (used for the finite-difference
check of the backward pass).
"""
from __future__ import annotations

import numpy as np

OFFS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def layer_shapes(cin=8, c1=16, c2=32, c3=64, cout=3):
    """(ci, co) of L0..L10."""
    return [(cin, c1), (c1, c1), (c1, c2), (c2, c2), (c2, c3), (c3, c3),
            (c3 + c2, c2), (c2, c2), (c2 + c1, c1), (c1, c1), (c1, cout)]


def num_params(shapes):
    return sum(27 * ci * co for ci, co in shapes)


def unflatten(params, shapes):
    out, q = [], 0
    for ci, co in shapes:
        m = 27 * ci * co
        out.append(np.asarray(params[q:q + m], dtype=np.float64).reshape(co, 27, ci))
        q += m
    return out


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _shift(x, off):
    """out[x] = in[x + off] (periodic) on (nb, nz, ny, nx, C)."""
    dz, dy, dx = off
    return np.roll(x, shift=(-dz, -dy, -dx), axis=(1, 2, 3))


def conv(x, w):
    co = w.shape[0]
    y = np.zeros(x.shape[:4] + (co,))
    for t, off in enumerate(OFFS):
        y += _shift(x, off) @ w[:, t, :].T
    return y


def conv_grad_input(dy, w):
    """dX[x, ci] = sum_{tap, co} W[co, tap, ci] dY[x - off, co]."""
    ci = w.shape[2]
    dx = np.zeros(dy.shape[:4] + (ci,))
    for t, off in enumerate(OFFS):
        dx += _shift(dy, tuple(-o for o in off)) @ w[:, t, :]
    return dx


def conv_grad_weight(dy, x):
    """dW[co, tap, ci] = sum_x dY[x, co] X[x + off, ci]."""
    co, ci = dy.shape[-1], x.shape[-1]
    dw = np.zeros((co, 27, ci))
    dyf = dy.reshape(-1, co)
    for t, off in enumerate(OFFS):
        dw[:, t, :] = dyf.T @ _shift(x, off).reshape(-1, ci)
    return dw


def pool(x):
    nb, nz, ny, nx, c = x.shape
    return x.reshape(nb, nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(2, 4, 6))


def pool_grad(dp):
    return upsample(dp) / 8.0


def upsample(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def upsample_grad(du):
    nb, nz, ny, nx, c = du.shape
    return du.reshape(nb, nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(2, 4, 6))


def relu(z):
    return np.maximum(z, 0.0)


def forward(params, x0, shapes, bf=False):
    """Returns y and the cache of stored activations."""
    R = bf16 if bf else (lambda a: a)
    W = [R(w) for w in unflatten(params, shapes)]
    c = {"W": W}
    x0 = R(x0)
    c["x0"] = x0
    c["a1"] = R(relu(conv(x0, W[0])))
    c["s1"] = R(relu(conv(c["a1"], W[1])))
    c["p1"] = R(pool(c["s1"]))
    c["a2"] = R(relu(conv(c["p1"], W[2])))
    c["s2"] = R(relu(conv(c["a2"], W[3])))
    c["p2"] = R(pool(c["s2"]))
    c["a3"] = R(relu(conv(c["p2"], W[4])))
    c["b3"] = R(relu(conv(c["a3"], W[5])))
    c["cat2"] = np.concatenate([upsample(c["b3"]), c["s2"]], axis=-1)
    c["d2"] = R(relu(conv(c["cat2"], W[6])))
    c["u2"] = R(relu(conv(c["d2"], W[7])))
    c["cat1"] = np.concatenate([upsample(c["u2"]), c["s1"]], axis=-1)
    c["d1"] = R(relu(conv(c["cat1"], W[8])))
    c["u1"] = R(relu(conv(c["d1"], W[9])))
    y = conv(c["u1"], W[10])  # kept in FP32 on the device, not rounded to bf16
    return y, c


def loss_and_grad(params, x0, target, shapes, bf=False):
    """Loss sum((y - t)^2) / (2 y.size) and its gradient in the flat parameter layout."""
    R = bf16 if bf else (lambda a: a)
    y, c = forward(params, x0, shapes, bf)
    W = c["W"]
    diff = y - target
    loss = float((diff ** 2).sum() / (2 * y.size))
    g = [None] * 11
    dy = R(diff / y.size)
    g[10] = conv_grad_weight(dy, c["u1"])

    def back(dout, out, layer, xin):
        dz = R(dout * (out > 0))
        g[layer] = conv_grad_weight(dz, xin)
        return dz

    du1 = R(conv_grad_input(dy, W[10]))
    dz = back(du1, c["u1"], 9, c["d1"])
    dd1 = R(conv_grad_input(dz, W[9]))
    dz = back(dd1, c["d1"], 8, c["cat1"])
    dcat1 = R(conv_grad_input(dz, W[8]))
    c2 = c["u2"].shape[-1]
    du2 = R(upsample_grad(dcat1[..., :c2]))
    ds1_skip = dcat1[..., c2:]
    dz = back(du2, c["u2"], 7, c["d2"])
    dd2 = R(conv_grad_input(dz, W[7]))
    dz = back(dd2, c["d2"], 6, c["cat2"])
    dcat2 = R(conv_grad_input(dz, W[6]))
    c3 = c["b3"].shape[-1]
    db3 = R(upsample_grad(dcat2[..., :c3]))
    ds2_skip = dcat2[..., c3:]
    dz = back(db3, c["b3"], 5, c["a3"])
    da3 = R(conv_grad_input(dz, W[5]))
    dz = back(da3, c["a3"], 4, c["p2"])
    dp2 = R(conv_grad_input(dz, W[4]))
    ds2 = R(ds2_skip + pool_grad(dp2))
    dz = back(ds2, c["s2"], 3, c["a2"])
    da2 = R(conv_grad_input(dz, W[3]))
    dz = back(da2, c["a2"], 2, c["p1"])
    dp1 = R(conv_grad_input(dz, W[2]))
    ds1 = R(ds1_skip + pool_grad(dp1))
    dz = back(ds1, c["s1"], 1, c["a1"])
    da1 = R(conv_grad_input(dz, W[1]))
    back(da1, c["a1"], 0, c["x0"])
    return loss, np.concatenate([gw.ravel() for gw in g])
