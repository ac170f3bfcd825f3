"""CPU restatement of the parameter selector's network (test infrastructure:
the checker of the library's MLP kernels, never the thing measured).

The network has no counterpart in the reference (its drivers use the analytic
estimates of PeriodicEwald/src/matlab/StokesSLP_ewald_2p.m:147-202).  Layout
of the packed parameters, as the library's kernels use them:
  W1[64][8], b1[64], W2[64][64], b2[64], W3[4][64], b3[4]
z1 = W1 x + b1, a1 = gelu(z1), z2 = W2 a1 + b2, a2 = gelu(z2), y = W3 a2 + b3,
gelu(z) = z Phi(z); loss = mean over samples and outputs of (y - t)^2.
Everything in float64.  Pinned by central finite differences
(tests/test_oracle_mlp.py)."""
import numpy as np
from scipy.special import erf

N_IN, N_HID, N_OUT = 8, 64, 4
SIZES = (("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
         ("W3", (N_OUT, N_HID)), ("b3", (N_OUT,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = w[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def forward(w, x):
    p = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ p["W1"].T + p["b1"]
    a1 = gelu(z1)
    z2 = a1 @ p["W2"].T + p["b2"]
    a2 = gelu(z2)
    return a2 @ p["W3"].T + p["b3"]


def loss_grad(w, x, t):
    """(loss, gradient) of mean((y - t)^2) by backpropagation."""
    p = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    z1 = x @ p["W1"].T + p["b1"]
    a1 = gelu(z1)
    z2 = a1 @ p["W2"].T + p["b2"]
    a2 = gelu(z2)
    y = a2 @ p["W3"].T + p["b3"]
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / r.size
    g = {"W3": dy.T @ a2, "b3": dy.sum(0)}
    d2 = (dy @ p["W3"]) * gelu_grad(z2)
    g["W2"] = d2.T @ a1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ p["W2"]) * gelu_grad(z1)
    g["W1"] = d1.T @ x
    g["b1"] = d1.sum(0)
    return loss, pack(g)


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba, 2015); returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
