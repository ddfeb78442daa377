"""FP64 restatement of the learned preconditioner's MLP (forward, loss,
backward) -- TEST INFRASTRUCTURE ONLY: the checker the device training
kernels are compared against.

Network (no counterpart in the reference, which uses plain CG in
include/iblgf/solver/linsys/linsys.hpp:170-230):
    A1 = tanh(X W1^T + b1), A2 = tanh(A1 W2^T + b2), Y = A2 W3^T + b3,
    loss = mean over rows and outputs of (Y - T)^2.
Parameters are one flat vector in the order W1 [H][D], b1 [H], W2 [H][H],
b2 [H], W3 [3][H], b3 [3].  The device stores W1, W2, W3 in BF16 (round to
nearest even); bf16_round reproduces that rounding.
"""
import numpy as np


def bf16_round(w):
    """Round FP32 values to BF16 (nearest even), returned as FP32."""
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(p, D, H):
    p = np.asarray(p)
    o = 0
    out = []
    for shape in ((H, D), (H,), (H, H), (H,), (3, H), (3,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    return out


def loss_and_grad(p, X, T, D, H, bf16_weights=True):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64), D, H)
    if bf16_weights:
        W1, W2, W3 = (bf16_round(W.astype(np.float32)).astype(np.float64) for W in (W1, W2, W3))
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    A1 = np.tanh(X @ W1.T + b1)
    A2 = np.tanh(A1 @ W2.T + b2)
    Y = A2 @ W3.T + b3
    n = Y.size
    loss = np.sum((Y - T) ** 2) / n
    dY = 2.0 * (Y - T) / n
    gW3 = dY.T @ A2
    gb3 = dY.sum(0)
    dZ2 = (dY @ W3) * (1 - A2 ** 2)
    gW2 = dZ2.T @ A1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (1 - A1 ** 2)
    gW1 = dZ1.T @ X
    gb1 = dZ1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def forward(p, X, D, H, bf16_weights=True):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64), D, H)
    if bf16_weights:
        W1, W2, W3 = (bf16_round(W.astype(np.float32)).astype(np.float64) for W in (W1, W2, W3))
    A1 = np.tanh(np.asarray(X, np.float64) @ W1.T + b1)
    A2 = np.tanh(A1 @ W2.T + b2)
    return A2 @ W3.T + b3
