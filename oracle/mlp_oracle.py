"""CPU restatement of the learned ID-rank predictor (test infrastructure).

TEST INFRASTRUCTURE ONLY: the GPU library never calls this module.  Only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import
it, and only as the checker.

The predictor has no counterpart in the reference: it is a 12-64-64-1
perceptron (tanh, tanh, sigmoid) mapping box features to the rank fraction
k / n that SkelFactorization::id_compress (skel_factorization.cpp:36-58)
and ki_Mat::id (ki_mat.cpp:525-568) produce for the box.  Its labels are
those ranks, so they are pinned by the ID oracle in ki_oracle.py.

Flat parameter layout (shared with the HIP kernels): W1 [64][12] (row =
hidden unit), b1 [64], W2 [64][64], b2 [64], W3 [64], b3.

Pinning of this module: the analytic gradient is checked against central
finite differences in FP64 (tests/test_oracle_mlp.py), the forward pass
against closed forms (zero weights give sigmoid(b3); a single active
unit gives a hand-computed value), and the Adam step against its
first-step closed form (|update| = lr for every nonzero gradient).
"""
from __future__ import annotations

import numpy as np

NF, NH = 12, 64
OW1 = 0
OB1 = OW1 + NH * NF
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
NP = OB3 + 1  # 5057


def unpack(p):
    p = np.asarray(p)
    return (p[OW1:OB1].reshape(NH, NF), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH),
            p[OB2:OW3], p[OW3:OB3], p[OB3])


def init_params(seed=0):
    """Glorot-uniform weights, zero biases (the same rule as ki_mlp_init;
    the random streams differ, so tests compare behaviour, not values)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NP)

    def glorot(fo, fi):
        a = np.sqrt(6.0 / (fi + fo))
        return rng.uniform(-a, a, fo * fi)
    p[OW1:OB1] = glorot(NH, NF)
    p[OW2:OB2] = glorot(NH, NH)
    p[OW3:OB3] = glorot(1, NH)
    return p


def forward(p, X):
    """Predicted rank fractions of the rows of X (n x 12)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    z = h2 @ W3 + b3
    return 1.0 / (1.0 + np.exp(-z))


def loss_and_grad(p, X, y):
    """Mean squared error of forward(p, X) against y and its gradient."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = len(y)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    yh = 1.0 / (1.0 + np.exp(-(h2 @ W3 + b3)))
    diff = yh - y
    loss = np.mean(diff ** 2)
    dz = 2.0 * diff * yh * (1.0 - yh) / n
    g = np.zeros(NP)
    g[OW3:OB3] = dz @ h2
    g[OB3] = dz.sum()
    da2 = np.outer(dz, W3) * (1.0 - h2 ** 2)
    g[OW2:OB2] = (da2.T @ h1).ravel()
    g[OB2:OW3] = da2.sum(0)
    da1 = (da2 @ W2) * (1.0 - h1 ** 2)
    g[OW1:OB1] = (da1.T @ X).ravel()
    g[OB1:OW2] = da1.sum(0)
    return loss, g


def finite_difference_grad(p, X, y, idx, h=1e-6):
    """Central differences of the loss for the parameters `idx`."""
    out = np.zeros(len(idx))
    for t, i in enumerate(idx):
        pp = p.copy()
        pp[i] += h
        lp = loss_and_grad(pp, X, y)[0]
        pp[i] -= 2 * h
        lm = loss_and_grad(pp, X, y)[0]
        out[t] = (lp - lm) / (2 * h)
    return out


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def box_features(level, active, original, side, m, special, tol, sol, dom,
                 points, normals):
    """The 12 features of a box (same definitions as the engine's):
    level/8, log2(n)/12, log2(n_original+1)/12, log2(m)/16, -log10(tol)/10,
    sol/3, dom/3, the active points' bounding-box extents / side in
    descending order, 1 - |mean unit normal|, and the special-row flag."""
    active = np.asarray(active)
    pts_idx = np.unique(active // sol)
    P = points[pts_idx]
    N = normals[pts_idx]
    ext = np.zeros(3)
    if len(P):
        ext[:dom] = (P.max(0) - P.min(0)) / side if side > 0 else 0.0
    ext = np.sort(ext)[::-1]
    mn = np.linalg.norm(N.sum(0)) / len(P) if len(P) else 1.0
    return np.array([level / 8.0, np.log2(max(len(active), 1)) / 12.0,
                     np.log2(len(original) + 1.0) / 12.0,
                     np.log2(max(m, 1)) / 16.0, -np.log10(tol) / 10.0,
                     sol / 3.0, dom / 3.0, ext[0], ext[1], ext[2], 1.0 - mn,
                     1.0 if special else 0.0])
