"""CPU restatement of the realignment screen's network (test infrastructure only: the tests compare
the HIP kernels' scores and gradients with it; nothing in the product imports it).

The network has no counterpart in the reference.  It follows the arithmetic the kernels in
rabbitvar_amd/csrc/rv_screen.hip document: BF16 operands (round to nearest even), sums in higher
precision here (FP32 in the kernels' MFMA accumulators), ReLU hidden layers whose outputs are
rounded to BF16, and the gradients of the mean weighted binary cross-entropy with the same BF16
roundings at the same points.  With bf=False every rounding is skipped and the arithmetic is
float64, which the tests check against central finite differences.
"""
import numpy as np

F, H = 32, 64
W1, B1 = 0, F * H
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H
NPARAM = B3 + 1


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(P):
    P = np.asarray(P, np.float64)
    return (P[W1:B1].reshape(F, H), P[B1:W2], P[W2:B2].reshape(H, H), P[B2:W3], P[W3:B3], P[B3])


def _r(x, bf):
    return bf16(x).astype(np.float64) if bf else np.asarray(x, np.float64)


def forward(P, X, bf=True):
    """Returns (y, cache): the logits and the intermediate values the backward pass reads."""
    w1, b1, w2, b2, w3, b3 = unpack(P)
    xb = _r(X, bf)
    w1b, w2b, w3b = _r(w1, bf), _r(w2, bf), _r(w3, bf)
    z1 = xb @ w1b + b1
    h1 = _r(np.maximum(z1, 0), bf)
    z2 = h1 @ w2b + b2
    h2 = _r(np.maximum(z2, 0), bf)
    y = h2 @ w3b + b3
    return y, (xb, h1, h2, w2b, w3b)


def scores(P, X, bf=True):
    y, _ = forward(P, X, bf)
    return 1.0 / (1.0 + np.exp(-y))


def loss_and_grad(P, X, t, pos_weight=1.0, bf=True):
    """Mean over the samples of w * BCE(sigmoid(y), t), w = pos_weight where t = 1, and its gradient
    with respect to the parameter vector (layout: W1[F][H], b1, W2[H][H], b2, w3, b3)."""
    t = np.asarray(t, np.float64)
    n = len(t)
    y, (xb, h1, h2, w2b, w3b) = forward(P, X, bf)
    w = np.where(t > 0, pos_weight, 1.0)
    z = np.where(t > 0, -y, y)
    loss = float(np.sum(w * (np.maximum(z, 0) + np.log1p(np.exp(-np.abs(z))))) / n)
    p = 1.0 / (1.0 + np.exp(-y))
    dy = w * (p - t) / n
    g = np.zeros(NPARAM)
    g[W3:B3] = h2.T @ dy
    g[B3] = dy.sum()
    g2 = _r(np.where(h2 > 0, dy[:, None] * w3b[None, :], 0.0), bf)
    g[W2:B2] = (h1.T @ g2).reshape(-1)
    g[B2:W3] = g2.sum(axis=0)
    g1 = np.where(h1 > 0, _r(g2 @ w2b.T, bf), 0.0)
    g[W1:B1] = (xb.T @ g1).reshape(-1)
    g[B1:W2] = g1.sum(axis=0)
    return loss, g
