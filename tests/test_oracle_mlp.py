"""CPU checks that pin the predictor oracle (oracle/mlp_oracle.py)."""
import numpy as np

import mlp_oracle as M


def _data(n=40, seed=1):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (n, M.NF))
    y = rng.uniform(0, 1, n)
    return X, y


def test_layout_sizes():
    assert M.NP == 64 * 12 + 64 + 64 * 64 + 64 + 64 + 1 == 5057


def test_gradient_matches_finite_differences():
    X, y = _data()
    p = M.init_params(3)
    p[M.OB1:M.OW2] = np.random.default_rng(4).normal(0, 0.3, M.NH)
    p[M.OB3] = 0.2
    loss, g = M.loss_and_grad(p, X, y)
    rng = np.random.default_rng(5)
    # a sample of every parameter block, biases included
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False)
                          for a, b in [(M.OW1, M.OB1), (M.OB1, M.OW2),
                                       (M.OW2, M.OB2), (M.OB2, M.OW3),
                                       (M.OW3, M.OB3)]] + [[M.OB3]])
    fd = M.finite_difference_grad(p, X, y, idx)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-9)
    assert loss == np.mean((M.forward(p, X) - y) ** 2)


def test_forward_zero_weights_is_sigmoid_of_output_bias():
    p = np.zeros(M.NP)
    p[M.OB3] = 0.7
    X, _ = _data(5)
    assert np.allclose(M.forward(p, X), 1 / (1 + np.exp(-0.7)))


def test_forward_single_path_layout():
    # only feature f -> hidden unit j -> hidden unit j2 -> output
    f, j, j2 = 3, 5, 9
    a, b, c = 0.8, -1.3, 2.1
    p = np.zeros(M.NP)
    p[M.OW1 + j * M.NF + f] = a
    p[M.OW2 + j2 * M.NH + j] = b
    p[M.OW3 + j2] = c
    X, _ = _data(6)
    want = 1 / (1 + np.exp(-c * np.tanh(b * np.tanh(a * X[:, f]))))
    assert np.allclose(M.forward(p, X), want)
    # the transposed placement couples a different feature/unit pair
    q = np.zeros(M.NP)
    q[M.OW1 + f * M.NF + j] = a
    q[M.OW2 + j2 * M.NH + f] = b
    q[M.OW3 + j2] = c
    want_t = 1 / (1 + np.exp(-c * np.tanh(b * np.tanh(a * X[:, j]))))
    assert np.allclose(M.forward(q, X), want_t)
    assert not np.allclose(want, want_t)


def test_adam_first_step_moves_each_parameter_by_lr():
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    p = np.ones(4)
    p1, m, v = M.adam_step(p, g, np.zeros(4), np.zeros(4), 1, 0.01)
    assert np.allclose((p - p1)[:3], 0.01 * np.sign(g[:3]), rtol=1e-4)
    assert p1[3] == 1.0


def test_features_of_a_flat_box():
    # four collinear points in 2D, box side 2: extents (1.5, 0, 0)/2
    pts = np.array([[0.0, 0.0], [0.5, 0.0], [1.0, 0.0], [1.5, 0.0]])
    nrm = np.tile([0.0, 1.0], (4, 1))
    f = M.box_features(level=3, active=[0, 1, 2, 3], original=[0, 1, 2, 3],
                       side=2.0, m=256, special=False, tol=1e-8, sol=1,
                       dom=2, points=pts, normals=nrm)
    assert np.allclose(f, [3 / 8, 2 / 12, np.log2(5) / 12, 8 / 16, 0.8,
                           1 / 3, 2 / 3, 0.75, 0.0, 0.0, 0.0, 0.0])
    # opposite normals: spread 1
    nrm2 = np.array([[0, 1.0], [0, -1.0], [0, 1.0], [0, -1.0]])
    f2 = M.box_features(3, [0, 1, 2, 3], [0, 1, 2, 3], 2.0, 256, True, 1e-8,
                        1, 2, pts, nrm2)
    assert f2[10] == 1.0 and f2[11] == 1.0
