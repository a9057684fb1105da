"""CPU checks of the learned-guess network's oracle (oracle/surrogate_oracle.py): the
exact float64 network's gradient against central finite differences, closed forms of
its building blocks, and the bfloat16 rounding rule."""
import numpy as np
import pytest

import surrogate_oracle as S


def he_params(rng, head_sd=None):
    lay, n = S.layout()
    th = np.zeros(n)
    for name, wo, nw, bo, ks, ci, co, relu in lay:
        sd = np.sqrt(2.0 / (ks * ks * ci)) if (head_sd is None or name != "head") else head_sd
        th[wo:wo + nw] = sd * rng.standard_normal(nw)
        th[bo:bo + co] = 0.05 * rng.standard_normal(co)
    return th


def test_parameter_count():
    lay, n = S.layout()
    assert n == sum(co * ks * ks * ci + co for ks, ci, co, _ in S.LAYERS)
    assert 0.9e6 < n < 1.2e6          # "~1M parameters"
    assert lay[-1][3] + lay[-1][6] == n


def test_bf16_rounding():
    one = 1.0
    cases = [(1.0, 1.0), (1 + 2.0 ** -8, 1.0), (1 + 3 * 2.0 ** -8, 1 + 2.0 ** -6),
             (1 + 2.0 ** -7, 1 + 2.0 ** -7), (-(1 + 2.0 ** -8 + 2.0 ** -12), -(1 + 2.0 ** -7)),
             (0.0, 0.0), (2.0 ** -130, 2.0 ** -130)]
    for x, want in cases:
        assert S.bf16(np.array([x]))[0] == want, (x, want)
    r = np.random.default_rng(0).standard_normal(1000)
    rel = np.abs(S.bf16(r) - r) / np.abs(r)
    assert rel.max() <= 2.0 ** -8 * (1 + 1e-12) * one


def test_conv_building_blocks():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 8, 8, 4))
    # a delta kernel at tap (ky, kx) reads the pixel (y + ky - 1, x + kx - 1), periodic
    for ky in range(3):
        for kx in range(3):
            W = np.zeros((4, 9 * 4))
            t = 3 * ky + kx
            W[:, t * 4:(t + 1) * 4] = np.eye(4)
            y = S.conv(x, W, np.zeros(4), 3, False)
            for (yy, xx) in [(0, 0), (7, 3), (4, 7)]:
                np.testing.assert_array_equal(y[1, yy, xx], x[1, (yy + ky - 1) % 8, (xx + kx - 1) % 8])
    # constant input: out = b + sum over taps of W (periodic padding has no border effect)
    W = rng.standard_normal((3, 9 * 4))
    b = rng.standard_normal(3)
    c = np.ones((1, 8, 8, 4)) * 0.5
    y = S.conv(c, W, b, 3, False)
    np.testing.assert_allclose(y, np.broadcast_to(b + 0.5 * W.sum(axis=1), y.shape), rtol=1e-13)
    # pooling / upsampling adjoints: <P a, b> = <a, P^T b>, <U a, b> = <a, U^T b>
    a, bb = rng.standard_normal((2, 8, 8, 3)), rng.standard_normal((2, 4, 4, 3))
    np.testing.assert_allclose((S.pool(a) * bb).sum(), (a * S.pool_backward(bb)).sum(), rtol=1e-13)
    np.testing.assert_allclose((S.up(bb) * a).sum(), (bb * S.up_backward(a)).sum(), rtol=1e-13)
    # conv adjoint in x: <conv(x), g> = <x, dX(g)> for zero bias
    W = rng.standard_normal((5, 9 * 4))
    g = rng.standard_normal((2, 8, 8, 5))
    _, _, dx = S.conv_backward(x, W, g, 3, False)
    np.testing.assert_allclose((S.conv(x, W, np.zeros(5), 3, False) * g).sum(), (x * dx).sum(), rtol=1e-12)


@pytest.mark.parametrize("nch", [24, 7])
def test_gradient_matches_finite_differences(nch):
    rng = np.random.default_rng(2)
    th = he_params(rng)
    x5 = rng.standard_normal((1, 8, 8, 5))
    tgt = rng.standard_normal((1, 8, 8, S.N_OUT))
    mask = (rng.random((1, 8, 8)) > 0.3).astype(float)
    L, g = S.loss_grad(th, x5, tgt, mask, nch=nch, bf16_ops=False)
    lay, n = S.layout()
    idx = []
    for name, wo, nw, bo, ks, ci, co, relu in lay:
        if name == "e0a":        # only the first 5 of 16 input channels carry data
            picks = [wo + c0 * ks * ks * ci + t * ci + c for c0, t, c in [(0, 4, 0), (3, 0, 4), (31, 8, 2)]]
        else:
            picks = list(wo + rng.integers(0, nw, 3))
        idx += picks + [bo + int(rng.integers(0, co))]
    # a channel that the loss does not see has zero gradient at the head
    eps = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += eps
        tm[i] -= eps
        Lp, _ = S.loss_grad(tp, x5, tgt, mask, nch=nch, bf16_ops=False)
        Lm, _ = S.loss_grad(tm, x5, tgt, mask, nch=nch, bf16_ops=False)
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])
    # zero-padded input channels 5..15 get exactly zero gradient
    name, wo, nw, bo, ks, ci, co, relu = lay[0]
    Wg = g[wo:wo + nw].reshape(co, ks * ks, ci)
    assert np.all(Wg[:, :, 5:] == 0.0)
    # head rows beyond nch get zero gradient
    name, wo, nw, bo, ks, ci, co, relu = lay[-1]
    assert np.all(g[wo:wo + nw].reshape(co, ci)[nch:] == 0.0) and np.all(g[bo + nch:bo + co] == 0.0)


def test_adam_first_step_is_sign_step():
    rng = np.random.default_rng(3)
    th = rng.standard_normal(50)
    g = rng.standard_normal(50)
    t1, m, v = S.adam(th, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    np.testing.assert_allclose(t1, th - 1e-3 * np.sign(g), rtol=0, atol=1e-7)
    # constant gradient: every later bias-corrected step is again lr * sign(g)
    t2, m, v = S.adam(t1, g, m, v, 2, 1e-3)
    np.testing.assert_allclose(t2, th - 2e-3 * np.sign(g), rtol=0, atol=1e-7)
