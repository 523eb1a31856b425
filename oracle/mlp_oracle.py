"""CPU oracle of the learned apsides-bracket predictor — TEST INFRASTRUCTURE ONLY.

The predictor has no counterpart in the reference: it proposes a starting bracket
for find_zero_Q (cmc_nr.c:316-383), and the exact bisection decides the result.
This module restates the device network in numpy so that the device forward and
backward kernels can be checked:

    x (6) -> standardise -> W1 (64x6) + b1 -> SiLU -> W2 (64x64) + b2 -> SiLU
          -> W3 (2x64) + b3 -> y (2), standardised (log r_peri, log r_apo)

    loss = mean over the batch of sum over the 2 outputs of (y - t)^2

Parameters are one flat float32 vector in the device's order:
    W1[64][6], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]   (4738 values)

Pinning: gradients against central finite differences in float64
(tests/test_oracle_mlp.py); SiLU and its derivative against closed forms.
"""
import numpy as np

NIN, NH, NOUT = 6, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NOUT * NH
NPARAM = OFF_B3 + NOUT


def unpack(theta):
    t = np.asarray(theta)
    return (t[OFF_W1:OFF_B1].reshape(NH, NIN), t[OFF_B1:OFF_W2], t[OFF_W2:OFF_B2].reshape(NH, NH),
            t[OFF_B2:OFF_W3], t[OFF_W3:OFF_B3].reshape(NOUT, NH), t[OFF_B3:NPARAM])


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def dsilu(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def forward(theta, x):
    """x: (B, 6) standardised inputs. Returns y (B, 2) and the cache."""
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    z1 = x @ W1.T + b1
    h1 = silu(z1)
    z2 = h1 @ W2.T + b2
    h2 = silu(z2)
    y = h2 @ W3.T + b3
    return y, (x, z1, h1, z2, h2)


def loss_and_grad(theta, x, t):
    """Mean squared error over the batch (summed over the two outputs) and its
    gradient with respect to the flat parameter vector."""
    y, (x, z1, h1, z2, h2) = forward(theta, x)
    W1, b1, W2, b2, W3, b3 = unpack(theta)
    B = x.shape[0]
    r = y - t
    loss = float(np.sum(r * r) / B)
    d3 = 2.0 * r / B                       # (B, 2)
    gW3 = d3.T @ h2
    gb3 = d3.sum(0)
    d2 = (d3 @ W3) * dsilu(z2)             # (B, 64)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * dsilu(z1)
    gW1 = d1.T @ x
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g


def init_params(seed):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for every weight and bias."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, n in ((NIN, NH * NIN), (NIN, NH), (NH, NH * NH), (NH, NH), (NH, NOUT * NH), (NH, NOUT)):
        a = 1.0 / np.sqrt(fan_in)
        parts.append(rng.uniform(-a, a, n))
    return np.concatenate(parts).astype(np.float32)


def adam_step(theta, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba 2015, Algorithm 1), t counted from 1."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
