"""FP64 oracle of the learned warm-start network -- test infrastructure only.

Restates, in float64 NumPy, what mechanic_amd/csrc/mlp_net.h and mlp.hip
compute: the 16 -> 64 -> 64 -> 3 network with GELU (tanh form,
gelu(x) = x * sigmoid(2 sqrt(2/pi) (x + 0.044715 x^3))), its input
normalisation and output scaling, the squared-error loss the GPU trainer
minimises and that loss's gradient. `forward(..., bf16=True)` also rounds
inputs, weights and hidden activations to BF16 at the points where the MFMA
inference does, so the GPU result can be checked to FP32-accumulation accuracy.

The network has no counterpart in the reference (its warm start is the halved
cached impulse, crates/mechanic-gpu/src/shaders/collision.wgsl:699-733, 783),
so nothing here cites it. Pinning, in tests/test_mlp_oracle.py: the GELU against
the textbook 0.5 x (1 + tanh(u)) form and the erf-based GELU, its derivative
and the whole loss gradient against central finite differences, and hand-built
networks whose output has a closed form (a single path through chosen units,
which a transposed weight matrix or swapped layer would break).

Only tests/, the smoke check and the CPU leg of the benchmark may use this file.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 16, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_H
OFF_W2 = OFF_B1 + N_H
OFF_B2 = OFF_W2 + N_H * N_H
OFF_W3 = OFF_B2 + N_H
OFF_B3 = OFF_W3 + N_H * N_OUT
N_PARAMS = OFF_B3 + N_OUT
OFF_IN_MEAN = N_PARAMS
OFF_IN_SCALE = OFF_IN_MEAN + N_IN
OFF_OUT_MEAN = OFF_IN_SCALE + N_IN
OFF_OUT_SCALE = OFF_OUT_MEAN + N_OUT
N_NET = OFF_OUT_SCALE + N_OUT

GELU_C = np.sqrt(2.0 / np.pi)
GELU_A = 0.044715


def unpack(net):
    net = np.asarray(net, np.float64)
    assert net.size == N_NET
    return dict(
        W1=net[OFF_W1:OFF_B1].reshape(N_IN, N_H), b1=net[OFF_B1:OFF_W2],
        W2=net[OFF_W2:OFF_B2].reshape(N_H, N_H), b2=net[OFF_B2:OFF_W3],
        W3=net[OFF_W3:OFF_B3].reshape(N_H, N_OUT), b3=net[OFF_B3:N_PARAMS],
        in_mean=net[OFF_IN_MEAN:OFF_IN_SCALE], in_scale=net[OFF_IN_SCALE:OFF_OUT_MEAN],
        out_mean=net[OFF_OUT_MEAN:OFF_OUT_SCALE], out_scale=net[OFF_OUT_SCALE:N_NET])


def gelu(x):
    x = np.asarray(x, np.float64)
    u = GELU_C * (x + GELU_A * x ** 3)
    with np.errstate(over="ignore"):  # exp overflows to inf for x << 0: gelu -> -0
        return x / (1.0 + np.exp(-2.0 * u))


def gelu_grad(x):
    x = np.asarray(x, np.float64)
    u = GELU_C * (x + GELU_A * x ** 3)
    with np.errstate(over="ignore"):
        s = 1.0 / (1.0 + np.exp(-2.0 * u))
    du = GELU_C * (1.0 + 3.0 * GELU_A * x * x)
    return s + x * s * (1.0 - s) * 2.0 * du


def bf16(x):
    """Round to the nearest BF16 (ties to even) via float32, as v_cvt_pk_bf16_f32 does."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(net, X, bf16_mode: bool = False):
    """Outputs (R, 3) in target units: (net(x_normalised)) * out_scale + out_mean."""
    p = unpack(net)
    X = np.asarray(X, np.float64)
    q = bf16 if bf16_mode else (lambda v: v)
    if bf16_mode:
        # The GPU normalises in FP32 before rounding to BF16.
        x32 = (X.astype(np.float32) - p["in_mean"].astype(np.float32)) * p["in_scale"].astype(np.float32)
        xn = bf16(x32)
    else:
        xn = (X - p["in_mean"]) * p["in_scale"]
    h1 = q(gelu(xn @ q(p["W1"]) + p["b1"]))
    h2 = q(gelu(h1 @ q(p["W2"]) + p["b2"]))
    y = h2 @ q(p["W3"]) + p["b3"]
    return y * p["out_scale"] + p["out_mean"]


def loss_and_grad(net, X, T):
    """Mean over rows of the summed squared error of the normalised outputs, and
    its gradient with respect to the N_PARAMS weights (float64)."""
    p = unpack(net)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    R = X.shape[0]
    xn = (X - p["in_mean"]) * p["in_scale"]
    tn = (T - p["out_mean"]) / p["out_scale"]
    z1 = xn @ p["W1"] + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = gelu(z2)
    y = h2 @ p["W3"] + p["b3"]
    err = y - tn
    loss = float(np.sum(err * err) / R)
    dy = 2.0 * err / R
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = (h2.T @ dy).ravel()
    g[OFF_B3:N_PARAMS] = dy.sum(0)
    dz2 = (dy @ p["W3"].T) * gelu_grad(z2)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = (dz2 @ p["W2"].T) * gelu_grad(z1)
    g[OFF_W1:OFF_B1] = (xn.T @ dz1).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g


def loss_only(net, X, T):
    p = unpack(net)
    xn = (np.asarray(X, np.float64) - p["in_mean"]) * p["in_scale"]
    tn = (np.asarray(T, np.float64) - p["out_mean"]) / p["out_scale"]
    y = gelu(gelu(xn @ p["W1"] + p["b1"]) @ p["W2"] + p["b2"]) @ p["W3"] + p["b3"]
    return float(np.sum((y - tn) ** 2) / X.shape[0])


def normalisation(features, targets):
    """The per-column statistics mh_mlp_set_dataset stores (float64 sums)."""
    f = np.asarray(features, np.float64)
    t = np.asarray(targets, np.float64)
    out = {}
    m = f.mean(0)
    var = (f * f).mean(0) - m * m
    out["in_mean"] = m
    out["in_scale"] = 1.0 / np.where(var > 1e-12, np.sqrt(np.maximum(var, 0.0)), 1.0)
    m = t.mean(0)
    var = (t * t).mean(0) - m * m
    out["out_mean"] = m
    out["out_scale"] = np.where(var > 1e-12, np.sqrt(np.maximum(var, 0.0)), 1.0)
    return out


def random_net(rng, scale: float = 1.0):
    """A random network with non-trivial normalisation, for tests."""
    net = np.zeros(N_NET)
    net[OFF_W1:OFF_B1] = rng.uniform(-1, 1, N_IN * N_H) * np.sqrt(6.0 / N_IN) * scale
    net[OFF_B1:OFF_W2] = rng.normal(0, 0.1, N_H)
    net[OFF_W2:OFF_B2] = rng.uniform(-1, 1, N_H * N_H) * np.sqrt(6.0 / N_H) * scale
    net[OFF_B2:OFF_W3] = rng.normal(0, 0.1, N_H)
    net[OFF_W3:OFF_B3] = rng.uniform(-1, 1, N_H * N_OUT) * np.sqrt(6.0 / N_H) * scale
    net[OFF_B3:N_PARAMS] = rng.normal(0, 0.1, N_OUT)
    net[OFF_IN_MEAN:OFF_IN_SCALE] = rng.normal(0, 1, N_IN)
    net[OFF_IN_SCALE:OFF_OUT_MEAN] = rng.uniform(0.5, 2.0, N_IN)
    net[OFF_OUT_MEAN:OFF_OUT_SCALE] = rng.normal(0, 1, N_OUT)
    net[OFF_OUT_SCALE:N_NET] = rng.uniform(0.5, 2.0, N_OUT)
    return net
