"""Pins of the oracle's proposal network (the checker of the GPU training
kernels), by what the mathematics fixes rather than by restating it:

* gradient = central finite differences of the loss, every parameter group;
* with zero biases a tanh network is an odd function (a wrong activation,
  e.g. a sigmoid or a ReLU, breaks this);
* a single active path W1[0][0] = W2[0][0] = W3[0] = 1 gives y = tanh(tanh(x0));
* zero weights give y = b3 and d(loss)/d(b3) = mean(y - t), loss = mean((y-t)^2)/2.
"""
import numpy as np

import oracle_lib as O

NW = 4545
W1, B1, W2, B2, W3, B3 = 0, 256, 320, 4416, 4480, 4544


def _params(rng, scale=0.3):
    P = np.zeros(NW + 10)
    P[:NW] = rng.normal(0., scale, NW)
    return P


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = _params(rng)
    X = rng.normal(size=(7, 4))
    T = rng.normal(size=7)
    G, loss = O.mlp_grad(P, X, T)
    Y = O.mlp_forward(P, X)
    assert abs(loss - np.mean(0.5 * (Y - T) ** 2)) < 1e-14
    idx = list(range(0, 256, 37)) + list(range(256, 320, 13)) + list(range(320, 4416, 311)) + \
        list(range(4416, 4480, 17)) + list(range(4480, 4544, 9)) + [4544]
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (O.mlp_grad(Pp, X, T)[1] - O.mlp_grad(Pm, X, T)[1]) / (2 * h)
        assert abs(fd - G[i]) <= 1e-7 * max(1.0, abs(G[i])), (i, fd, G[i])


def test_odd_function_without_biases():
    rng = np.random.default_rng(4)
    P = _params(rng)
    for lo, hi in ((B1, W2), (B2, W3), (B3, B3 + 1)):
        P[lo:hi] = 0.
    X = rng.normal(size=(20, 4))
    np.testing.assert_allclose(O.mlp_forward(P, -X), -O.mlp_forward(P, X), rtol=0, atol=1e-15)


def test_single_path_closed_form():
    P = np.zeros(NW + 10)
    P[W1 + 0] = 1.0      # W1[0][0]
    P[W2 + 0] = 1.0      # W2[0][0]
    P[W3 + 0] = 1.0      # W3[0]
    X = np.zeros((5, 4))
    X[:, 0] = np.linspace(-2, 2, 5)
    X[:, 1:] = 7.0       # other inputs do not reach the output
    np.testing.assert_allclose(O.mlp_forward(P, X), np.tanh(np.tanh(X[:, 0])), rtol=0, atol=1e-15)


def test_zero_weights_constant_output():
    P = np.zeros(NW + 10)
    P[B3] = 0.7
    rng = np.random.default_rng(5)
    X = rng.normal(size=(9, 4))
    T = rng.normal(size=9)
    assert np.all(O.mlp_forward(P, X) == 0.7)
    G, loss = O.mlp_grad(P, X, T)
    assert abs(G[B3] - np.mean(0.7 - T)) < 1e-15
    assert abs(loss - np.mean(0.5 * (0.7 - T) ** 2)) < 1e-15
    assert np.all(G[W2:B2] == 0.)  # h1 = 0 exactly
