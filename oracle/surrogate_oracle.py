"""CPU oracle of the learned-guess network (numpy).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the GPU network's
forward pass, parameter gradients and Adam step.  The network has no
counterpart in the reference, so this oracle restates the network's own
definition (the layer table, periodic 3x3 convolutions, 2x2 average pooling,
nearest upsampling, concatenation, masked MSE):

    enc0  3x3 16->32, 32->32            dec2  up(mid), cat enc2, 3x3 256->128
    enc1  pool, 3x3 32->64, 64->64      dec1  up(dec2), cat enc1, 3x3 192->64
    enc2  pool, 3x3 64->128, 128->128   dec0  up(dec1), cat enc0, 3x3 96->32
    mid   pool, 3x3 128->128 x2         head  1x1 32->32 (no ReLU)

Parameter vector: per layer W[co][tap*Ci + ci] (tap = 3*ky + kx reads pixel
(y + ky - 1, x + kx - 1), periodic), then b[co]; layers in that order.

With bf16=True every convolution operand (input activation, weight, and in the
backward pass the output gradient) is rounded to bfloat16 (round to nearest
even) before an exact (float64) contraction, which is what the MFMA kernels
compute up to FP32 accumulation order; bias gradients and everything outside
the contractions stay unrounded.  With bf16=False it is the exact float64
network, pinned by central finite differences (tests/test_surrogate_oracle.py).
"""
from __future__ import annotations

import numpy as np

N_IN_RAW = 5
N_IN = 16
N_OUT = 24
N_OUT_PAD = 32

# (ks, Ci, Co, relu)
LAYERS = [
    (3, N_IN, 32, 1), (3, 32, 32, 1), (3, 32, 64, 1), (3, 64, 64, 1),
    (3, 64, 128, 1), (3, 128, 128, 1), (3, 128, 128, 1), (3, 128, 128, 1),
    (3, 256, 128, 1), (3, 192, 64, 1), (3, 96, 32, 1), (1, 32, N_OUT_PAD, 0),
]
NAMES = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "m0", "m1", "d2", "d1", "d0", "head"]


def layout():
    """[(name, woff, nw, boff, ks, Ci, Co, relu)] and the total parameter count."""
    out, off = [], 0
    for name, (ks, ci, co, relu) in zip(NAMES, LAYERS):
        nw = co * ks * ks * ci
        out.append((name, off, nw, off + nw, ks, ci, co, relu))
        off += nw + co
    return out, off


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _taps(ks):
    if ks == 1:
        return [(0, 0)]
    return [(ky - 1, kx - 1) for ky in range(3) for kx in range(3)]


def _shift(x, dy, dx):
    """out[b, y, x] = x[b, y + dy, x + dx] (periodic)."""
    return np.roll(x, shift=(-dy, -dx), axis=(1, 2))


def conv(x, W, b, ks, round_ops):
    """x (B, H, W, Ci); W (Co, ks*ks*Ci); b (Co,) -> (B, H, W, Co)."""
    ci = x.shape[3]
    xr = bf16(x) if round_ops else x
    Wr = bf16(W) if round_ops else W
    out = np.zeros(x.shape[:3] + (W.shape[0],))
    for t, (dy, dx) in enumerate(_taps(ks)):
        out += _shift(xr, dy, dx) @ Wr[:, t * ci:(t + 1) * ci].T
    return out + b


def conv_backward(x, W, dy_out, ks, round_ops, want_dx=True):
    ci = x.shape[3]
    xr = bf16(x) if round_ops else x
    Wr = bf16(W) if round_ops else W
    gr = bf16(dy_out) if round_ops else dy_out
    co = W.shape[0]
    g2 = gr.reshape(-1, co)
    dW = np.zeros_like(W, dtype=np.float64)
    dx_ = np.zeros(x.shape) if want_dx else None
    for t, (dy, dx) in enumerate(_taps(ks)):
        xs = _shift(xr, dy, dx).reshape(-1, ci)
        dW[:, t * ci:(t + 1) * ci] = g2.T @ xs
        if want_dx:
            dx_ += _shift(gr @ Wr[:, t * ci:(t + 1) * ci], -dy, -dx)
    db = dy_out.reshape(-1, co).sum(axis=0)
    return dW, db, dx_


def pool(x):
    B, H, W, C = x.shape
    return x.reshape(B, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_backward(d):
    return 0.25 * up(d)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_backward(d):
    B, H, W, C = d.shape
    return d.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def widen(x5):
    x = np.zeros(x5.shape[:3] + (N_IN,))
    x[..., :N_IN_RAW] = x5
    return x


def _params(theta):
    lay, n = layout()
    assert theta.size == n, (theta.size, n)
    P = {}
    for name, wo, nw, bo, ks, ci, co, relu in lay:
        P[name] = (theta[wo:wo + nw].reshape(co, ks * ks * ci).astype(np.float64),
                   theta[bo:bo + co].astype(np.float64), ks, relu)
    return P


def forward(theta, x5, bf16_ops=True, cache=None):
    """Network output (B, H, W, 32) for input (B, H, W, 5)."""
    P = _params(theta)
    a = {}

    def c(name, inp):
        W, b, ks, relu = P[name]
        y = conv(inp, W, b, ks, bf16_ops)
        if bf16_ops:
            y = y.astype(np.float32).astype(np.float64)   # activations are stored as float32
        return np.maximum(y, 0.0) if relu else y

    a["in"] = widen(x5)
    a["e0a"] = c("e0a", a["in"]); a["e0b"] = c("e0b", a["e0a"]); a["p0"] = pool(a["e0b"])
    a["e1a"] = c("e1a", a["p0"]); a["e1b"] = c("e1b", a["e1a"]); a["p1"] = pool(a["e1b"])
    a["e2a"] = c("e2a", a["p1"]); a["e2b"] = c("e2b", a["e2a"]); a["p2"] = pool(a["e2b"])
    a["m0"] = c("m0", a["p2"]); a["m1"] = c("m1", a["m0"])
    a["c2"] = np.concatenate([up(a["m1"]), a["e2b"]], axis=3); a["d2"] = c("d2", a["c2"])
    a["c1"] = np.concatenate([up(a["d2"]), a["e1b"]], axis=3); a["d1"] = c("d1", a["c1"])
    a["c0"] = np.concatenate([up(a["d1"]), a["e0b"]], axis=3); a["d0"] = c("d0", a["c0"])
    a["out"] = c("head", a["d0"])
    if cache is not None:
        cache.update(a)
    return a["out"]


def loss_of(out, tgt, mask, nch, chw=None):
    chw = np.ones(N_OUT) if chw is None else np.asarray(chw, dtype=np.float64)
    m = mask.sum()
    e = out[..., :nch] - tgt[..., :nch]
    L = float((mask[..., None] * chw[:nch] * e * e).sum() / m)
    dout = np.zeros_like(out)
    dout[..., :nch] = 2.0 * mask[..., None] * chw[:nch] * e / m
    return L, dout


def loss_grad(theta, x5, tgt, mask, nch=N_OUT, bf16_ops=True, chw=None):
    """(loss, gradient vector) of the masked MSE."""
    P = _params(theta)
    a = {}
    out = forward(theta, x5, bf16_ops, cache=a)
    L, dout = loss_of(out, tgt, mask, nch, chw)
    lay, n = layout()
    off = {name: (wo, nw, bo, co) for name, wo, nw, bo, ks, ci, co, relu in lay}
    grad = np.zeros(n)

    def back(name, inp, outp, dy, want_dx=True):
        W, b, ks, relu = P[name]
        if relu:
            dy = np.where(outp > 0, dy, 0.0)
        dW, db, dx = conv_backward(inp, W, dy, ks, bf16_ops, want_dx)
        wo, nw, bo, co = off[name]
        grad[wo:wo + nw] = dW.ravel()
        grad[bo:bo + co] = db
        return dx

    def split(d, ca):
        return up_backward(d[..., :ca]), d[..., ca:]

    dd0 = back("head", a["d0"], a["out"], dout)
    dc0 = back("d0", a["c0"], a["d0"], dd0)
    dd1, de0b = split(dc0, a["d1"].shape[3])
    dc1 = back("d1", a["c1"], a["d1"], dd1)
    dd2, de1b = split(dc1, a["d2"].shape[3])
    dc2 = back("d2", a["c2"], a["d2"], dd2)
    dm1, de2b = split(dc2, a["m1"].shape[3])
    dm0 = back("m1", a["m0"], a["m1"], dm1)
    dp2 = back("m0", a["p2"], a["m0"], dm0)
    de2b = de2b + pool_backward(dp2)
    de2a = back("e2b", a["e2a"], a["e2b"], de2b)
    dp1 = back("e2a", a["p1"], a["e2a"], de2a)
    de1b = de1b + pool_backward(dp1)
    de1a = back("e1b", a["e1a"], a["e1b"], de1b)
    dp0 = back("e1a", a["p0"], a["e1a"], de1a)
    de0b = de0b + pool_backward(dp0)
    de0a = back("e0b", a["e0a"], a["e0b"], de0b)
    back("e0a", a["in"], a["e0a"], de0a, want_dx=False)
    return L, grad


def adam(theta, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (step counts from 1); returns new (theta, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    th = theta - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return th, m, v
