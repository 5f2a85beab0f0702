"""NumPy oracle for the surrogate MLP — test infrastructure only.

f(x) = w3 . silu(W2 silu(W1 x + b1) + b2) + b3 on normalised inputs; the
input gradient g = df/dx; the training loss per sample
    l = alpha (f - y)^2 + beta/D ||g - t||^2
and its gradients with respect to all weights, written out by hand (double
backpropagation through g).  The reference has no surrogate; this restates
the mathematics above and is pinned in tests by central finite differences
on the loss and on f.
"""
from __future__ import annotations

import numpy as np


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def d2silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    ds = s * (1.0 - s)
    return ds * (2.0 + z * (1.0 - 2.0 * s))


def forward(P, X):
    """X (B, D) -> f (B,), g (B, D) and the intermediates."""
    W1, b1, W2, b2, w3, b3 = P["W1"], P["b1"], P["W2"], P["b2"], P["w3"], P["b3"]
    Z1 = X @ W1.T + b1
    A1 = silu(Z1)
    Z2 = A1 @ W2.T + b2
    A2 = silu(Z2)
    f = A2 @ w3 + b3
    S1, S2 = dsilu(Z1), dsilu(Z2)
    U2 = S2 * w3
    V1 = U2 @ W2
    U1 = S1 * V1
    G = U1 @ W1
    return f, G, dict(Z1=Z1, A1=A1, Z2=Z2, A2=A2, S1=S1, S2=S2, U2=U2, V1=V1, U1=U1)


def loss_and_grads(P, X, y, T, alpha=1.0, beta=1.0):
    """Mean loss over the batch and its gradients (same keys as P)."""
    B, D = X.shape
    f, G, c = forward(P, X)
    W1, W2, w3 = P["W1"], P["W2"], P["w3"]
    r = f - y
    E = G - T
    loss = np.mean(alpha * r * r + beta / D * np.sum(E * E, axis=1))
    df = 2.0 * alpha * r / B                       # dl/df       (B,)
    dG = 2.0 * beta / D * E / B                    # dl/dg       (B, D)
    # g = U1 W1
    dW1 = c["U1"].T @ dG                           # (H, D)
    dU1 = dG @ W1.T                                # (B, H)
    # U1 = S1 * V1
    dS1 = dU1 * c["V1"]
    dV1 = dU1 * c["S1"]
    # V1 = U2 W2
    dW2 = c["U2"].T @ dV1                          # (H, H): d/dW2[j, i] of sum_j U2[j] W2[j, i]
    dU2 = dV1 @ W2.T
    # U2 = S2 * w3
    dS2 = dU2 * w3
    dw3 = np.sum(dU2 * c["S2"], axis=0)
    # f = A2 w3 + b3
    dw3 += c["A2"].T @ df
    db3 = np.sum(df)
    dA2 = np.outer(df, w3)
    # S2 = silu'(Z2), A2 = silu(Z2)
    dZ2 = dS2 * d2silu(c["Z2"]) + dA2 * c["S2"]
    dW2 += dZ2.T @ c["A1"]
    db2 = np.sum(dZ2, axis=0)
    dA1 = dZ2 @ W2
    # S1 = silu'(Z1), A1 = silu(Z1)
    dZ1 = dS1 * d2silu(c["Z1"]) + dA1 * c["S1"]
    dW1 += dZ1.T @ X
    db1 = np.sum(dZ1, axis=0)
    return loss, dict(W1=dW1, b1=db1, W2=dW2, b2=db2, w3=dw3, b3=db3)


def adam_step(P, grads, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Adam (Kingma & Ba 2015) with bias correction; updates in place."""
    for k in P:
        m[k] = b1 * m[k] + (1 - b1) * grads[k]
        v[k] = b2 * v[k] + (1 - b2) * grads[k] ** 2
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        P[k] = P[k] - lr * mh / (np.sqrt(vh) + eps)
    return P
