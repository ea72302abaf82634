"""ORACLE — test infrastructure only.

FP64 numpy restatement of the direction network of libachr
(achr.cu_amd/csrc/achr_dirnet.hip): F(z) = W3 gelu(W2 gelu(W1 z + b1) + b2) + b3,
v(z) = F(z) - F(-z), loss |V^T V / B - C|_F^2 / |C|_F^2, with the gradient
by reverse-mode differentiation written out by hand. The gradient is pinned
by central finite differences (tests/test_dirnet.py); the GPU gradient is
checked against it.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

SQRT2 = np.sqrt(2.0)


def unpack(p, L, H, O):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * L].reshape(H, L); o += H * L
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + O * H].reshape(O, H); o += O * H
    b3 = p[o:o + O]; o += O
    return W1, b1, W2, b2, W3, b3


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / SQRT2))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / SQRT2)) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def forward(p, Z, L, H, O):
    W1, b1, W2, b2, W3, b3 = unpack(p, L, H, O)
    Z2 = np.vstack([Z, -Z]).astype(np.float64)
    A1 = Z2 @ W1.T + b1
    H1 = gelu(A1)
    A2 = H1 @ W2.T + b2
    H2 = gelu(A2)
    F = H2 @ W3.T + b3
    B = Z.shape[0]
    V = F[:B] - F[B:]
    return V, (Z2, A1, H1, A2, H2)


def loss(p, Z, C, L, H, O):
    V, _ = forward(p, Z, L, H, O)
    B = Z.shape[0]
    E = V.T @ V / B - C
    return float((E * E).sum() / (C * C).sum())


def loss_grad(p, Z, C, L, H, O):
    W1, b1, W2, b2, W3, b3 = unpack(p, L, H, O)
    V, (Z2, A1, H1, A2, H2) = forward(p, Z, L, H, O)
    B = Z.shape[0]
    E = V.T @ V / B - C
    c2 = (C * C).sum()
    dV = 4.0 / (B * c2) * V @ E
    dF = np.vstack([dV, -dV])
    gW3 = dF.T @ H2
    gb3 = dF.sum(0)
    dA2 = (dF @ W3) * gelu_grad(A2)
    gW2 = dA2.T @ H1
    gb2 = dA2.sum(0)
    dA1 = (dA2 @ W2) * gelu_grad(A1)
    gW1 = dA1.T @ Z2
    gb1 = dA1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return float((E * E).sum() / c2), g
