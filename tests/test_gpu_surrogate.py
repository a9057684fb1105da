"""The learned-guess network on the GPU (MFMA BF16 convolutions, hand-written backward,
Adam) against its CPU oracle (oracle/surrogate_oracle.py, BF16 operands emulated)."""
import numpy as np
import pytest

import surrogate_oracle as S
from cism_amd.surrogate import Surrogate

pytestmark = pytest.mark.gpu

B, H, W = 2, 16, 24


def he_params(rng):
    lay, n = S.layout()
    th = np.zeros(n, dtype=np.float32)
    for name, wo, nw, bo, ks, ci, co, relu in lay:
        th[wo:wo + nw] = np.sqrt(2.0 / (ks * ks * ci)) * rng.standard_normal(nw)
        th[bo:bo + co] = 0.05 * rng.standard_normal(co)
    return th


def batch(rng):
    x5 = rng.standard_normal((B, H, W, 5)).astype(np.float32)
    tgt = rng.standard_normal((B, H, W, S.N_OUT)).astype(np.float32)
    mask = (rng.random((B, H, W)) > 0.25).astype(np.float32)
    return x5, tgt, mask


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def net():
    print(f"[surrogate] create {B}x{H}x{W}", flush=True)
    n = Surrogate(H, W, B, 0)
    yield n
    n.close()


def test_param_count(net):
    assert net.nparams == S.layout()[1]


def test_forward_matches_oracle(net):
    rng = np.random.default_rng(10)
    th = he_params(rng)
    x5, _, _ = batch(rng)
    net.set_params(th)
    print("[surrogate] forward", flush=True)
    g = net.forward(x5)
    o = S.forward(th.astype(np.float64), x5.astype(np.float64), bf16_ops=True)[..., :S.N_OUT]
    e = rel(g, o)
    print(f"[surrogate] forward rel L2 vs oracle {e:.2e}")
    assert e < 5e-3
    # the exact float64 network differs by the BF16 operand rounding only
    o64 = S.forward(th.astype(np.float64), x5.astype(np.float64), bf16_ops=False)[..., :S.N_OUT]
    assert rel(g, o64) < 5e-2


@pytest.mark.parametrize("nch", [24, 10])
def test_gradients_match_oracle(net, nch):
    rng = np.random.default_rng(11 + nch)
    th = he_params(rng)
    x5, tgt, mask = batch(rng)
    net.set_params(th)
    print(f"[surrogate] loss_grad nch={nch}", flush=True)
    Lg, gg = net.loss_grad(x5, tgt, mask, nch)
    Lo, go = S.loss_grad(th.astype(np.float64), x5.astype(np.float64), tgt.astype(np.float64),
                         mask.astype(np.float64), nch=nch, bf16_ops=True)
    assert abs(Lg - Lo) <= 1e-3 * Lo, (Lg, Lo)
    lay, _ = S.layout()
    for name, wo, nw, bo, ks, ci, co, relu in lay:
        ew = rel(gg[wo:wo + nw], go[wo:wo + nw])
        eb = rel(gg[bo:bo + co], go[bo:bo + co])
        print(f"[surrogate] grad {name}: W rel {ew:.2e}, b rel {eb:.2e}")
        assert ew < 2e-2 and eb < 2e-2, (name, ew, eb)
    # the zero-padded input channels and unseen head rows get exactly zero gradient
    name, wo, nw, bo, ks, ci, co, relu = lay[0]
    assert np.all(gg[wo:wo + nw].reshape(co, 9, ci)[:, :, 5:] == 0)
    name, wo, nw, bo, ks, ci, co, relu = lay[-1]
    assert np.all(gg[wo:wo + nw].reshape(co, ci)[nch:] == 0)


def test_adam_step_matches_oracle(net):
    rng = np.random.default_rng(20)
    th = he_params(rng)
    x5, tgt, mask = batch(rng)
    net.set_params(th)
    _, g = net.loss_grad(x5, tgt, mask)
    net.set_params(th)                       # resets the moments
    net.train_step(x5, tgt, mask, 1e-3)
    t1 = net.get_params()
    want, _, _ = S.adam(th.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    np.testing.assert_allclose(t1, want, rtol=0, atol=2e-6)


def test_training_reduces_loss(net):
    rng = np.random.default_rng(30)
    th = he_params(rng)
    lay, _ = S.layout()
    name, wo, nw, bo, ks, ci, co, relu = lay[-1]
    th[wo:wo + nw] *= 0.1
    x5, _, mask = batch(rng)
    # a smooth target that is a function of the input: learnable
    tgt = np.zeros((B, H, W, S.N_OUT), np.float32)
    tgt[..., :5] = np.tanh(x5)
    tgt[..., 5:10] = 0.5 * x5 ** 2 - 0.5
    net.set_params(th)
    losses = [net.train_step(x5, tgt, mask, 1e-3, nch=10) for _ in range(120)]
    print(f"[surrogate] training loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    # the same 120 steps run through the CPU oracle (float64 with BF16 operands) take the
    # loss from 4.71 to 0.66
    assert losses[-1] < 0.3 * losses[0]
