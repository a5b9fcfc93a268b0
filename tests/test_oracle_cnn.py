"""Pin the oracle of the learned coarse-grid correction network on the CPU.

The backward pass of oracle/cnn.py is checked against central finite
differences of its own FP64 loss (directional derivatives along random
directions and single weights in every layer); the building blocks against
direct definitions (a convolution against a brute-force triple loop, the
adjointness <conv(x), y> = <x, conv_grad_input(y)>, pooling/upsampling
adjoint pairs).  A wrong tap offset, a transposed weight or a dropped skip
path fails one of these.
"""
import numpy as np
import pytest

from oracle import cnn as N


def _rand(shape, seed, scale=1.0):
    return scale * np.random.default_rng(seed).standard_normal(shape)


def _params(shapes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for ci, co in shapes:
        out.append(rng.standard_normal(27 * ci * co) * np.sqrt(2.0 / (27 * ci)))
    return np.concatenate(out)


def test_conv_matches_brute_force():
    x = _rand((1, 4, 4, 6, 3), 1)
    w = _rand((5, 27, 3), 2)
    y = N.conv(x, w)
    nb, nz, ny, nx, _ = x.shape
    want = np.zeros((1, nz, ny, nx, 5))
    for z in range(nz):
        for yy in range(ny):
            for xx in range(nx):
                for t, (dz, dy, dx) in enumerate(N.OFFS):
                    v = x[0, (z + dz) % nz, (yy + dy) % ny, (xx + dx) % nx]
                    want[0, z, yy, xx] += w[:, t, :] @ v
    assert np.allclose(y, want, rtol=1e-13, atol=1e-13)


def test_conv_adjoints():
    x = _rand((2, 4, 6, 4, 3), 3)
    w = _rand((5, 27, 3), 4)
    dy = _rand((2, 4, 6, 4, 5), 5)
    lhs = (N.conv(x, w) * dy).sum()
    assert np.isclose(lhs, (x * N.conv_grad_input(dy, w)).sum(), rtol=1e-12)
    assert np.isclose(lhs, (w * N.conv_grad_weight(dy, x)).sum(), rtol=1e-12)
    p = _rand((1, 4, 4, 4, 2), 6)
    q = _rand((1, 2, 2, 2, 2), 7)
    assert np.isclose((N.pool(p) * q).sum(), (p * N.pool_grad(q)).sum(), rtol=1e-12)
    assert np.isclose((N.upsample(q) * p).sum(), (q * N.upsample_grad(p)).sum(), rtol=1e-12)


def _small():
    shapes = N.layer_shapes(cin=4, c1=4, c2=4, c3=8, cout=3)
    p = _params(shapes, 11)
    x = _rand((2, 8, 8, 8, 4), 12)
    t = _rand((2, 8, 8, 8, 3), 13, 0.3)
    return shapes, p, x, t


def test_gradient_directional_finite_differences():
    shapes, p, x, t = _small()
    _, g = N.loss_and_grad(p, x, t, shapes)
    rng = np.random.default_rng(14)
    for _ in range(3):
        v = rng.standard_normal(p.size)
        eps = 1e-6
        lp, _ = N.loss_and_grad(p + eps * v, x, t, shapes)
        lm, _ = N.loss_and_grad(p - eps * v, x, t, shapes)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g @ v) <= 1e-6 * abs(fd) + 1e-12


def test_gradient_single_weights_every_layer():
    shapes, p, x, t = _small()
    _, g = N.loss_and_grad(p, x, t, shapes)
    rng = np.random.default_rng(15)
    q0 = 0
    for ci, co in shapes:
        m = 27 * ci * co
        for q in q0 + rng.choice(m, 3, replace=False):
            e = np.zeros(p.size)
            e[q] = 1.0
            eps = 1e-6
            lp, _ = N.loss_and_grad(p + eps * e, x, t, shapes)
            lm, _ = N.loss_and_grad(p - eps * e, x, t, shapes)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[q]) <= 1e-5 * max(abs(fd), abs(g[q])) + 1e-10
        q0 += m


def test_bf16_rounding():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3])
    r = N.bf16(a)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie goes to even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 2 ** -8 * 3.0e-3  # half an ulp of 8 significant bits
