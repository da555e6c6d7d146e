"""ORACLE — test infrastructure only (never imported by the product path).

numpy restatement (float64) of the warm-start MLP of pyqcm_amd/csrc/qed_mlp.hip: forward
pass, mean-squared-error loss and its gradient by backpropagation, and one Adam step.
The network has no counterpart in the reference (it is new in this solver); the
restatement is pinned by central finite differences of the loss (tests/test_oracle.py),
which a wrong transpose, a dropped tanh' factor or a wrong mean normalization fails.

Parameter layout (flat): W1 [H][n_in], b1 [H], W2 [H][H], b2 [H], W3 [n_out][H], b3.
"""
import numpy as np


def unpack(p, n_in, n_h, n_out):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + n_h * n_in].reshape(n_h, n_in); o += n_h * n_in
    b1 = p[o:o + n_h]; o += n_h
    W2 = p[o:o + n_h * n_h].reshape(n_h, n_h); o += n_h * n_h
    b2 = p[o:o + n_h]; o += n_h
    W3 = p[o:o + n_out * n_h].reshape(n_out, n_h); o += n_out * n_h
    b3 = p[o:o + n_out]; o += n_out
    assert o == len(p)
    return W1, b1, W2, b2, W3, b3


def forward(p, x, n_in, n_h, n_out):
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, n_h, n_out)
    h1 = np.tanh(x @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3.T + b3, (h1, h2)


def loss_grad(p, x, t, n_in, n_h, n_out):
    """loss = mean((y - t)^2) over samples and outputs; returns (loss, flat gradient)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, n_in, n_h, n_out)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    y, (h1, h2) = forward(p, x, n_in, n_h, n_out)
    inv = 1.0 / (len(x) * n_out)
    d = y - t
    loss = np.sum(d * d) * inv
    dy = 2.0 * d * inv
    dz2 = (dy @ W3) * (1.0 - h2 * h2)
    dz1 = (dz2 @ W2) * (1.0 - h1 * h1)
    g = np.concatenate([(dz1.T @ x).ravel(), dz1.sum(0), (dz2.T @ h1).ravel(), dz2.sum(0),
                        (dy.T @ h2).ravel(), dy.sum(0)])
    return loss, g


def adam_step(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
