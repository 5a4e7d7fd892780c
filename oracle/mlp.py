"""CPU restatement of the learned screener's network (test infrastructure only).

Forward and backward of the n_in-128-128-1 ReLU perceptron in FP64 numpy with
the same BF16 rounding of the weight matrices (round to nearest even) that the
GPU kernels apply, the weighted mean squared error L = sum_r w_r (y_r - t_r)^2
/ (2 n), and one Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias
corrected).  Parameter layout: W1 [128 x n_in], b1, W2 [128 x 128], b2,
W3 [128], b3.  The gradient is pinned by central finite differences on the
FP64 network in tests/test_mlp_oracle.py.
"""
import numpy as np

NH = 128


def bf16(x):
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p, nin):
    o = 0
    W1 = p[o:o + NH * nin].reshape(NH, nin); o += NH * nin
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    W3 = p[o:o + NH]; o += NH
    b3 = p[o:o + 1]; o += 1
    return W1, b1, W2, b2, W3, b3


def num_params(nin):
    return NH * nin + NH + NH * NH + NH + NH + 1


def forward(p, X, nin, round_bf16=True):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, dtype=np.float64), nin)
    if round_bf16:
        W1, W2, W3 = bf16(W1), bf16(W2), bf16(W3)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ W3 + b3[0]
    return y, (X, h1, h2, W1, W2, W3)


def loss_grad(p, X, t, w, nin, round_bf16=True):
    y, (X, h1, h2, W1, W2, W3) = forward(p, X, nin, round_bf16)
    n = len(y)
    w = np.ones(n) if w is None else np.asarray(w, dtype=np.float64)
    e = y - t
    loss = 0.5 * np.sum(w * e * e) / n
    dy = w * e / n
    gW3 = h2.T @ dy
    gb3 = np.array([dy.sum()])
    dz2 = np.outer(dy, W3) * (h2 > 0)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2 @ W2) * (h1 > 0)
    gW1 = dz1.T @ X
    gb1 = dz1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, gb3])
    return loss, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
