"""Pins the FP64 network oracle (oracle/mlp_oracle.py) by means independent of
its code path: the textbook tanh form of GELU and the erf-based GELU it
approximates, central finite differences for every derivative, and hand-built
networks whose outputs have closed forms."""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_gelu_matches_tanh_form_and_erf_gelu():
    xs = np.linspace(-8.0, 8.0, 4001)
    tanh_form = np.array([0.5 * x * (1.0 + math.tanh(math.sqrt(2.0 / math.pi) * (x + 0.044715 * x ** 3)))
                          for x in xs])
    np.testing.assert_allclose(mo.gelu(xs), tanh_form, rtol=1e-13, atol=1e-15)
    erf_form = np.array([0.5 * x * (1.0 + math.erf(x / math.sqrt(2.0))) for x in xs])
    # The tanh form approximates the erf GELU to a few 1e-4 everywhere.
    assert np.max(np.abs(mo.gelu(xs) - erf_form)) < 1e-3
    assert mo.gelu(0.0) == 0.0
    assert abs(mo.gelu(30.0) - 30.0) < 1e-12 and abs(mo.gelu(-30.0)) < 1e-12


def test_gelu_grad_central_difference():
    xs = np.linspace(-6.0, 6.0, 241)
    h = 1e-6
    fd = (mo.gelu(xs + h) - mo.gelu(xs - h)) / (2 * h)
    np.testing.assert_allclose(mo.gelu_grad(xs), fd, rtol=1e-7, atol=1e-8)


def test_bf16_rounding_ties_to_even():
    # 1 + 2^-8 is halfway between BF16 neighbours 1 and 1 + 2^-7: ties to even -> 1.
    assert mo.bf16(1.0 + 2.0 ** -8) == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6.
    assert mo.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert mo.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert mo.bf16(-2.5) == -2.5


def _identity_norm(net):
    net[mo.OFF_IN_MEAN:mo.OFF_IN_SCALE] = 0.0
    net[mo.OFF_IN_SCALE:mo.OFF_OUT_MEAN] = 1.0
    net[mo.OFF_OUT_MEAN:mo.OFF_OUT_SCALE] = 0.0
    net[mo.OFF_OUT_SCALE:mo.N_NET] = 1.0
    return net


def test_single_path_network_closed_form():
    """Input 3 -> hidden-1 unit 5 -> hidden-2 unit 41 -> output 1, every other
    weight zero: output 1 = w3 * gelu(w2 * gelu(w1 * x3)); a transposed W2 or
    W1, or swapped layers, routes the signal to a zero weight instead."""
    net = _identity_norm(np.zeros(mo.N_NET))
    w1, w2, w3 = 1.5, -0.75, 2.0
    net[mo.OFF_W1 + 3 * mo.N_H + 5] = w1        # W1[3][5]
    net[mo.OFF_W2 + 5 * mo.N_H + 41] = w2       # W2[5][41]
    net[mo.OFF_W3 + 41 * mo.N_OUT + 1] = w3     # W3[41][1]
    net[mo.OFF_B3:mo.N_PARAMS] = [0.25, -0.5, 0.125]
    X = np.zeros((7, mo.N_IN))
    X[:, 3] = np.linspace(-2, 2, 7)
    X[:, 0] = 9.0  # unused input
    y = mo.forward(net, X)

    def g(v):
        return 0.5 * v * (1.0 + math.tanh(math.sqrt(2.0 / math.pi) * (v + 0.044715 * v ** 3)))

    expect1 = np.array([w3 * g(w2 * g(w1 * x)) for x in X[:, 3]]) - 0.5
    np.testing.assert_allclose(y[:, 1], expect1, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(y[:, 0], 0.25)
    np.testing.assert_allclose(y[:, 2], 0.125)
    # The transposed W2 entry carries nothing.
    net2 = _identity_norm(np.zeros(mo.N_NET))
    net2[mo.OFF_W1 + 3 * mo.N_H + 5] = w1
    net2[mo.OFF_W2 + 41 * mo.N_H + 5] = w2
    net2[mo.OFF_W3 + 41 * mo.N_OUT + 1] = w3
    assert np.all(mo.forward(net2, X)[:, 1] == 0.0)


def test_normalisation_and_scaling():
    """x -> (x - mean) * scale on the way in, y * std + mean on the way out."""
    net = _identity_norm(np.zeros(mo.N_NET))
    net[mo.OFF_W1 + 2 * mo.N_H + 0] = 1.0     # W1[2][0]
    net[mo.OFF_W2 + 0 * mo.N_H + 0] = 1.0     # W2[0][0]
    net[mo.OFF_W3 + 0 * mo.N_OUT + 2] = 1.0   # W3[0][2]
    net[mo.OFF_IN_MEAN + 2] = 1.0
    net[mo.OFF_IN_SCALE + 2] = 4.0
    net[mo.OFF_OUT_MEAN + 2] = 10.0
    net[mo.OFF_OUT_SCALE + 2] = 3.0
    X = np.zeros((1, mo.N_IN))
    X[0, 2] = 1.5
    y = mo.forward(net, X)
    assert abs(y[0, 2] - (3.0 * mo.gelu(mo.gelu(2.0)) + 10.0)) < 1e-12
    assert y[0, 0] == 0.0 and y[0, 1] == 0.0


def test_loss_gradient_central_difference():
    rng = np.random.default_rng(3)
    net = mo.random_net(rng)
    X = rng.normal(0, 1.5, (40, mo.N_IN))
    T = rng.normal(0, 1.0, (40, mo.N_OUT))
    loss, g = mo.loss_and_grad(net, X, T)
    assert abs(loss - mo.loss_only(net, X, T)) < 1e-12 * max(1.0, loss)
    # Every bias and a spread of weights in each matrix.
    idx = list(range(mo.OFF_B1, mo.OFF_W2)) + list(range(mo.OFF_B2, mo.OFF_W3)) + \
        list(range(mo.OFF_B3, mo.N_PARAMS)) + list(rng.choice(np.arange(mo.OFF_W1, mo.OFF_B1), 60, replace=False)) + \
        list(rng.choice(np.arange(mo.OFF_W2, mo.OFF_B2), 120, replace=False)) + \
        list(rng.choice(np.arange(mo.OFF_W3, mo.OFF_B3), 60, replace=False))
    h = 1e-6
    for i in idx:
        a = net.copy()
        b = net.copy()
        a[i] += h
        b[i] -= h
        fd = (mo.loss_only(a, X, T) - mo.loss_only(b, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (1.0 + abs(g[i])), (i, fd, g[i])


def test_loss_is_a_mean_over_rows():
    rng = np.random.default_rng(5)
    net = mo.random_net(rng)
    X = rng.normal(0, 1, (30, mo.N_IN))
    T = rng.normal(0, 1, (30, mo.N_OUT))
    l_all, g_all = mo.loss_and_grad(net, X, T)
    l_a, g_a = mo.loss_and_grad(net, X[:10], T[:10])
    l_b, g_b = mo.loss_and_grad(net, X[10:], T[10:])
    assert abs(l_all - (10 * l_a + 20 * l_b) / 30) < 1e-12
    np.testing.assert_allclose(g_all, (10 * g_a + 20 * g_b) / 30, rtol=1e-10, atol=1e-13)
    perm = rng.permutation(30)
    l_p, g_p = mo.loss_and_grad(net, X[perm], T[perm])
    assert abs(l_p - l_all) < 1e-12
    np.testing.assert_allclose(g_p, g_all, rtol=1e-10, atol=1e-13)


def test_normalisation_statistics():
    rng = np.random.default_rng(9)
    f = rng.normal(3.0, 2.0, (500, mo.N_IN))
    f[:, 7] = 1.25  # constant column: unit scale
    t = rng.normal(-1.0, 0.5, (500, mo.N_OUT))
    s = mo.normalisation(f, t)
    np.testing.assert_allclose(s["in_mean"], f.mean(0), rtol=1e-12)
    np.testing.assert_allclose(1.0 / s["in_scale"][np.arange(16) != 7], f.std(0)[np.arange(16) != 7], rtol=1e-9)
    assert s["in_scale"][7] == 1.0
    np.testing.assert_allclose(s["out_scale"], t.std(0), rtol=1e-9)
