"""CPU oracle of the property-table MLP surrogate (TEST INFRASTRUCTURE ONLY).

Forward pass and hand-written backward pass of the mean squared error for
the 1 -> 64 -> 64 -> 5 tanh network, in float64 on weights given as arrays
(the product evaluates with BF16-rounded weights: pass those).  No
counterpart in the reference; pinned by central finite differences
(tests/test_oracle_mlp.py).  Parameter layout: W1[64], b1[64], W2[64][64]
(row = output unit), b2[64], W3[5][64], b3[5].
"""
import numpy as np

H1, H2, NO = 64, 64, 5
OW1, OB1 = 0, 64
OW2 = 128
OB2 = OW2 + H2 * H1
OW3 = OB2 + H2
OB3 = OW3 + NO * H2
NP = OB3 + NO


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OW1:OB1], p[OB1:OW2], p[OW2:OB2].reshape(H2, H1), p[OB2:OW3], p[OW3:OB3].reshape(NO, H2), p[OB3:NP])


def forward(p, x):
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    a1 = np.tanh(np.outer(x, w1) + b1)
    a2 = np.tanh(a1 @ w2.T + b2)
    return a2 @ w3.T + b3, a1, a2


def loss_grad(p, x, t):
    """mean over samples and outputs of (y - t)^2, and its gradient"""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    y, a1, a2 = forward(p, x)
    B = x.size
    e = y - t
    loss = np.mean(e**2)
    dy = 2.0 * e / (B * NO)
    g = np.zeros(NP)
    g[OW3:OB3] = (dy.T @ a2).ravel()
    g[OB3:NP] = dy.sum(axis=0)
    dz2 = (dy @ w3) * (1.0 - a2**2)
    g[OW2:OB2] = (dz2.T @ a1).ravel()
    g[OB2:OW3] = dz2.sum(axis=0)
    dz1 = (dz2 @ w2) * (1.0 - a1**2)
    g[OW1:OB1] = dz1.T @ x
    g[OB1:OW2] = dz1.sum(axis=0)
    return loss, g


def bf16_round(p):
    """round-to-nearest-even to BF16, returned as float64"""
    u = np.asarray(p, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
