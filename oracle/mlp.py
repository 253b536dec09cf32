"""CPU oracle of the rank-prediction network (test infrastructure only).

Float64 restatement of the forward pass and the back-propagated gradient of the mean
squared error for the MLP 8 -> 64 -> 64 -> 1 with the tanh-approximated GELU, using the
same flat parameter layout as hssc_mlp_get_params: W1 (8 x 64, column-major), b1, W2 (64 x
64), b2, w3 (64), b3.  Also restates the closed-form rank guess of HSS2D/predict_rank.m:5-23
that the network replaces.
"""
from __future__ import annotations

import math

import numpy as np

NIN, NH = 8, 64


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NIN * NH].reshape(NH, NIN).T
    o += NIN * NH
    b1 = p[o:o + NH]
    o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH).T
    o += NH * NH
    b2 = p[o:o + NH]
    o += NH
    w3 = p[o:o + NH]
    o += NH
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([W1.T.ravel(), b1, W2.T.ravel(), b2, w3, [b3]])


def gelu(z):
    s = math.sqrt(2 / math.pi)
    return 0.5 * z * (1 + np.tanh(s * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    s = math.sqrt(2 / math.pi)
    t = np.tanh(s * (z + 0.044715 * z ** 3))
    return 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * s * (1 + 3 * 0.044715 * z * z)


def standardise(X, mu, sd):
    return (np.asarray(X, np.float64) - mu[:NIN]) / sd[:NIN]


def forward(p, Xs):
    """Xs: B x 8 standardised features -> standardised outputs (B,), cache."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    Z1 = Xs @ W1 + b1
    A1 = gelu(Z1)
    Z2 = A1 @ W2 + b2
    A2 = gelu(Z2)
    y = A2 @ w3 + b3
    return y, (Xs, Z1, A1, Z2, A2)


def loss_and_grad(p, Xs, ts):
    """Mean squared error against standardised targets ts and its gradient (flat)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (X, Z1, A1, Z2, A2) = forward(p, Xs)
    B = X.shape[0]
    r = y - ts
    loss = float(np.mean(r * r))
    d = 2 * r / B
    gw3 = A2.T @ d
    gb3 = d.sum()
    D2 = np.outer(d, w3) * gelu_grad(Z2)
    gW2 = A1.T @ D2
    gb2 = D2.sum(0)
    D1 = (D2 @ W2.T) * gelu_grad(Z1)
    gW1 = X.T @ D1
    gb1 = D1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gw3, gb3)


def predict_rank(n, acc):
    """HSS2D/predict_rank.m:5-23: (rank, C) from the log-law tables."""
    p = int(math.ceil(-math.log10(acc)))
    p = min(12, max(7, p))
    m = {7: 96, 8: 138, 9: 184, 10: 275, 11: 526, 12: 487}[p]
    b = {7: 138, 8: 215, 9: 300, 10: 496, 11: 1074, 12: 954}[p]
    r0 = {7: 86, 8: 104, 9: 116, 10: 130, 11: 137, 12: 140}[p]
    C = int(math.ceil(m * math.log10(2) / 2))
    if n < 150:
        return min(n, r0), C
    return min(n, int(math.ceil(m * math.log10(n) - b))), C
