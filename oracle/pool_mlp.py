"""ORACLE — test infrastructure only.

binary64 numpy restatement of the learned pool-size network that the device
kernels train (``csrc/pool_mlp.hip``): inputs -> 64 tanh -> 64 tanh ->
outputs, parameters flat as [W1 (64 x I) | b1 | W2 (64 x 64) | b2 |
W3 (O x 64) | b3] (row-major), asymmetric squared error

    L = 1/(N O) sum_{s,o} w(r) r^2,  r = y - t,  w = under if r < 0 else 1,

its analytic gradient by reverse accumulation, and the Adam step (Kingma and
Ba, 2015). Nothing here is shipped; the tests check the device kernels against
these functions, and pin these functions by central finite differences.
"""

import numpy as np

HIDDEN = 64


def param_count(inputs, outputs):
    return HIDDEN * inputs + 2 * HIDDEN + HIDDEN * HIDDEN + outputs * HIDDEN + outputs


def unpack(params, inputs, outputs):
    p = np.asarray(params, dtype=np.float64)
    h = HIDDEN
    k = 0
    W1 = p[k:k + h * inputs].reshape(h, inputs); k += h * inputs
    b1 = p[k:k + h]; k += h
    W2 = p[k:k + h * h].reshape(h, h); k += h * h
    b2 = p[k:k + h]; k += h
    W3 = p[k:k + outputs * h].reshape(outputs, h); k += outputs * h
    b3 = p[k:k + outputs]; k += outputs
    assert k == len(p)
    return W1, b1, W2, b2, W3, b3


def forward(params, x, outputs):
    x = np.asarray(x, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(params, x.shape[1], outputs)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3.T + b3


def loss(params, x, t, under):
    t = np.asarray(t, dtype=np.float64)
    r = forward(params, x, t.shape[1]) - t
    w = np.where(r < 0, under, 1.0)
    return float(np.sum(w * r * r) / r.size)


def gradient(params, x, t, under):
    """(loss, dL/dparams) by reverse accumulation."""
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n, outputs = t.shape
    W1, b1, W2, b2, W3, b3 = unpack(params, x.shape[1], outputs)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    r = h2 @ W3.T + b3 - t
    w = np.where(r < 0, under, 1.0)
    value = float(np.sum(w * r * r) / r.size)
    dy = 2.0 * w * r / r.size
    dz2 = (dy @ W3) * (1.0 - h2 * h2)
    dz1 = (dz2 @ W2) * (1.0 - h1 * h1)
    grad = np.concatenate([
        (dz1.T @ x).ravel(), dz1.sum(0),
        (dz2.T @ h1).ravel(), dz2.sum(0),
        (dy.T @ h2).ravel(), dy.sum(0),
    ])
    return value, grad


def adam(params, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step at iteration ``step`` (1-based); returns (p, m, v)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mhat = m / (1 - beta1 ** step)
    vhat = v / (1 - beta2 ** step)
    return params - lr * mhat / (np.sqrt(vhat) + eps), m, v
