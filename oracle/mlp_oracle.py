"""FP64 numpy restatement of the warm start's network (TEST INFRASTRUCTURE
ONLY: the checker of the HIP kernels in pincosimulator_amd/csrc/twmc_mlp.hip,
used by tests/ and nothing else).

No reference counterpart.  The network is a fully connected MLP with
exact-erf GELU on the hidden layers and a linear output; parameters are one
flat vector holding, per layer, W [in][out] row-major then b [out]; the loss is
sum_ij w_ij (y_ij - t_ij)^2 / rows.  The backward pass is pinned by central
finite differences in tests/test_mlp_oracle.py, GELU by its closed-form
values and symmetry.
"""
import math

import numpy as np

_erf = np.vectorize(math.erf, otypes=[np.float64])


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def widths(n_in, n_hidden, n_hidden_layers, n_out):
    return [n_in] + [n_hidden] * n_hidden_layers + [n_out]


def unpack(params, ws):
    out, o = [], 0
    for a, b in zip(ws[:-1], ws[1:]):
        W = params[o:o + a * b].reshape(a, b)
        o += a * b
        bias = params[o:o + b]
        o += b
        out.append((W, bias))
    assert o == params.size
    return out


def forward(params, ws, x):
    """Returns (y, cache)."""
    layers = unpack(np.asarray(params, np.float64), ws)
    a = np.asarray(x, np.float64)
    acts, pres = [a], []
    for l, (W, b) in enumerate(layers):
        z = a @ W + b
        if l < len(layers) - 1:
            pres.append(z)
            a = gelu(z)
        else:
            a = z
        acts.append(a)
    return a, (acts, pres)


def loss_grad(params, ws, x, t, w=None):
    """(loss, flat gradient) of sum w (y - t)^2 / rows."""
    y, (acts, pres) = forward(params, ws, x)
    rows = y.shape[0]
    w = np.ones_like(y) if w is None else np.asarray(w, np.float64)
    d = y - np.asarray(t, np.float64)
    loss = float((w * d * d).sum() / rows)
    dz = 2.0 * w * d / rows
    layers = unpack(np.asarray(params, np.float64), ws)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (acts[l].T @ dz, dz.sum(axis=0))
        if l > 0:
            dz = (dz @ W.T) * gelu_grad(pres[l - 1])
    flat = np.concatenate([np.concatenate([g[0].reshape(-1), g[1]]) for g in grads])
    return loss, flat
