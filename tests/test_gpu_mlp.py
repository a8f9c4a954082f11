"""The warm start's network: HIP forward, backward and Adam vs the FP64 numpy
oracle (oracle/mlp_oracle.py, itself pinned by finite differences).

Bar (FP32 kernels vs FP64 oracle on the same FP32 parameters): outputs within
1e-5 relative to the output scale, each layer's weight and bias gradient
within 1e-4 relative to that layer's largest gradient entry, one Adam step
within 1e-6 absolute of the oracle's update (lr = 1e-3, eps = 0.1 so that the
step is a smooth function of the gradient and FP32 gradient rounding cannot
flip it)."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo
from pincosimulator_amd.mlp import MLP

pytestmark = pytest.mark.gpu

WS = (5, 34, 256, 3)  # n_in, n_out, hidden width, hidden layers


def net(seed=3):
    return MLP(WS[0], WS[1], n_hidden=WS[2], n_hidden_layers=WS[3], seed=seed)


def widths():
    return mo.widths(WS[0], WS[2], WS[3], WS[1])


def layer_slices():
    out, o = [], 0
    ws = widths()
    for a, b in zip(ws[:-1], ws[1:]):
        out.append((slice(o, o + a * b), slice(o + a * b, o + a * b + b)))
        o += a * b + b
    return out


def test_forward_matches_oracle():
    m = net()
    rng = np.random.default_rng(0)
    x = rng.normal(size=(77, 5)).astype(np.float32)  # not a multiple of the 64-row tile
    print("mlp forward", flush=True)
    y = m.forward(x)
    yo, _ = mo.forward(m.params.astype(np.float64), widths(), x.astype(np.float64))
    assert np.max(np.abs(y - yo)) < 1e-5 * np.max(np.abs(yo))
    m.close()


@pytest.mark.parametrize("rows", [64, 131])
def test_gradients_match_oracle(rows):
    m = net(seed=11)
    rng = np.random.default_rng(rows)
    x = rng.normal(size=(rows, 5)).astype(np.float32)
    t = rng.normal(size=(rows, 34)).astype(np.float32)
    w = (rng.uniform(size=(rows, 34)) > 0.2).astype(np.float32)
    print(f"mlp gradients rows={rows}", flush=True)
    loss, g = m.loss_grad(x, t, w)
    lo, go = mo.loss_grad(m.params.astype(np.float64), widths(), x.astype(np.float64), t.astype(np.float64),
                          w.astype(np.float64))
    assert abs(loss - lo) < 1e-5 * lo
    for sw, sb in layer_slices():
        for s in (sw, sb):
            scale = np.max(np.abs(go[s]))
            assert np.max(np.abs(g[s] - go[s])) < 1e-4 * scale
    m.close()


def test_one_adam_step_matches_oracle():
    m = net(seed=5)
    rng = np.random.default_rng(9)
    x = rng.normal(size=(128, 5)).astype(np.float32)
    t = rng.normal(size=(128, 34)).astype(np.float32)
    p0 = m.params.astype(np.float64)
    _, g = mo.loss_grad(p0, widths(), x.astype(np.float64), t.astype(np.float64))
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 0.1
    mm = (1 - b1) * g
    vv = (1 - b2) * g * g
    p1 = p0 - lr * (mm / (1 - b1)) / (np.sqrt(vv / (1 - b2)) + eps)
    print("mlp adam step", flush=True)
    m.train(x, t, epochs=1, batch=128, lr=lr, beta1=b1, beta2=b2, eps=eps)
    assert np.max(np.abs(m.params - p1)) < 1e-6
    m.close()


def test_training_fits_and_is_repeatable():
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, size=(2048, 5)).astype(np.float32)
    A = rng.normal(size=(5, 34)) * 0.8
    t = np.sin(x @ A + rng.uniform(0, 3, 34)).astype(np.float32)  # smooth targets
    print("mlp training", flush=True)
    a = net(seed=21)
    h = a.train(x, t, epochs=100, batch=256, lr=1e-3, shuffle_seed=4)
    print("epoch losses", h[0], h[9], h[-1])
    assert h[-1] < 0.05 * h[0], (h[0], h[-1])
    b = net(seed=21)
    b.train(x, t, epochs=100, batch=256, lr=1e-3, shuffle_seed=4)
    np.testing.assert_array_equal(a.params, b.params)
    a.close()
    b.close()
