"""Pins of the learned-preconditioner network restatement (numpy, FP64):
finite-difference gradients, the adjoint identity of conv, bfloat16 rounding."""
import numpy as np
import pytest

from oracle import lpc_oracle as lo


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -3.1415926, 0.0, 65504.0], dtype=np.float32)
    r = lo.bf16(x)
    # 8 significand bits: 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0 + 2 ** -6)
    assert r[3] == np.float32(-3.140625) and r[4] == 0.0
    assert np.all(np.abs(r - x) <= np.abs(x) * 2 ** -8)


def test_conv_adjoint_and_identity():
    rng = np.random.default_rng(0)
    W = rng.standard_normal((3, 2, 3, 3, 3))
    x = rng.standard_normal((2, 5, 4, 6))
    y = rng.standard_normal((3, 5, 4, 6))
    assert np.sum(lo.conv(W, x) * y) == pytest.approx(np.sum(x * lo.conv_t(W, y)), rel=1e-12)
    # centre tap only: a pointwise channel mix
    Wc = np.zeros((2, 2, 3, 3, 3))
    Wc[:, :, 1, 1, 1] = [[1, 2], [3, 4]]
    np.testing.assert_allclose(lo.conv(Wc, x), np.einsum("oc,c...->o...", Wc[:, :, 1, 1, 1], x))
    # one off-centre tap is a periodic shift: W[.., d3=2, d2=1, d1=0] reads x(i3+1, i2, i1-1)
    Ws = np.zeros((1, 1, 3, 3, 3))
    Ws[0, 0, 2, 1, 0] = 1.0
    np.testing.assert_allclose(lo.conv(Ws, x[:1])[0], np.roll(x[0], shift=(-1, 0, 1), axis=(0, 1, 2)))


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    n = (5, 4, 6)
    g = rng.standard_normal(n)
    v = rng.standard_normal(n)
    t = rng.standard_normal(n)
    p = lo.init_params(3)
    p["W3"] = rng.standard_normal(p["W3"].shape) * 0.3
    p["b1"] = rng.standard_normal(16) * 0.1
    L, gr = lo.loss_and_grads(p, g, v, t)
    h = 1e-6
    for name, idx in [("W1", (3, 1, 0, 2, 1)), ("W1", (7, 0, 1, 1, 1)), ("W2", (2, 5, 1, 0, 2)),
                      ("W3", (0, 9, 2, 2, 0)), ("b1", (4,)), ("b2", (11,)), ("b3", (0,))]:
        q = {k: v_.copy() for k, v_ in p.items()}
        q[name][idx] += h
        Lp, _ = lo.loss_and_grads(q, g, v, t)
        q[name][idx] -= 2 * h
        Lm, _ = lo.loss_and_grads(q, g, v, t)
        assert gr[name][idx] == pytest.approx((Lp - Lm) / (2 * h), rel=1e-5, abs=1e-9), (name, idx)


def test_output_is_scale_equivariant():
    rng = np.random.default_rng(2)
    n = (4, 4, 4)
    g, v = rng.standard_normal(n), rng.standard_normal(n)
    p = lo.init_params(0)
    a, _ = lo.forward(p, g, v)
    b, _ = lo.forward(p, 7.5 * g, v)
    np.testing.assert_allclose(b, 7.5 * a, rtol=1e-12)
