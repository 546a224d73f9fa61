"""FP64 numpy restatement of the controller network (test infrastructure only;
what libetd's etd_mlp forward / backward kernels are checked against).

MLP 16 -> 64 -> 64 -> 1, GELU(a) = a/2 (1 + erf(a / sqrt 2)), loss = mean
squared error.  Parameter vector layout as in include/etd.h: W1 [64][16]
row-major, b1 [64], W2 [64][64], b2 [64], w3 [64], b3.  The gradient is
pinned by central finite differences (tests/test_oracle.py)."""
import numpy as np
from scipy.special import erf

NF, NH = 16, 64
SIZES = [("W1", NH * NF), ("b1", NH), ("W2", NH * NH), ("b2", NH), ("w3", NH), ("b3", 1)]
NPAR = sum(s for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, s in SIZES:
        out[name] = p[o:o + s]
        o += s
    out["W1"] = out["W1"].reshape(NH, NF)
    out["W2"] = out["W2"].reshape(NH, NH)
    return out


def gelu(a):
    return 0.5 * a * (1.0 + erf(a / np.sqrt(2.0)))


def gelu_d(a):
    return 0.5 * (1.0 + erf(a / np.sqrt(2.0))) + a * np.exp(-0.5 * a * a) / np.sqrt(2.0 * np.pi)


def forward(p, x):
    q = unpack(p)
    x = np.asarray(x, dtype=np.float64).reshape(-1, NF)
    a1 = x @ q["W1"].T + q["b1"]
    h1 = gelu(a1)
    a2 = h1 @ q["W2"].T + q["b2"]
    h2 = gelu(a2)
    return h2 @ q["w3"] + q["b3"][0], (x, a1, h1, a2, h2)


def loss_grad(p, x, t):
    q = unpack(p)
    y, (x, a1, h1, a2, h2) = forward(p, x)
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    B = x.shape[0]
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / B
    g_w3 = h2.T @ dy
    g_b3 = np.array([dy.sum()])
    da2 = (dy[:, None] * q["w3"][None, :]) * gelu_d(a2)
    g_W2 = da2.T @ h1
    g_b2 = da2.sum(0)
    da1 = (da2 @ q["W2"]) * gelu_d(a1)
    g_W1 = da1.T @ x
    g_b1 = da1.sum(0)
    g = np.concatenate([g_W1.reshape(-1), g_b1, g_W2.reshape(-1), g_b2, g_w3, g_b3])
    return loss, g


def init(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPAR)
    q = unpack(p)
    l1, l2, l3 = np.sqrt(6.0 / (NF + NH)), np.sqrt(6.0 / (2 * NH)), np.sqrt(6.0 / (NH + 1))
    parts = [rng.uniform(-l1, l1, NH * NF), np.zeros(NH), rng.uniform(-l2, l2, NH * NH), np.zeros(NH),
             rng.uniform(-l3, l3, NH), np.zeros(1)]
    return np.concatenate(parts)
