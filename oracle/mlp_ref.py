"""TEST INFRASTRUCTURE ONLY: CPU (numpy, float64) restatement of the warm-
start MLP's forward pass and parameter gradient, used as the checker of the
device kernels (imex_sflow2d_v2_amd/csrc/imexsw_mlp.h).

The device rounds to bfloat16 at fixed points -- the inputs, the weights,
every hidden activation and every back-propagated pre-activation gradient --
and accumulates in float32.  `quantise=True` applies the same rounding at the
same points (round to nearest even), so the two differ only by the
accumulation order and precision; `quantise=False` is the exact float64
network, checked against central finite differences.
"""
import numpy as np

IN, IN_PAD, H, OUT, OUT_PAD = 9, 16, 64, 2, 16
OFF = {}
_o = 0
for _name, _n in (("W1", H * IN_PAD), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * H),
                  ("b3", H), ("W4", OUT_PAD * H), ("b4", OUT_PAD)):
    OFF[_name] = (_o, _n)
    _o += _n
N_PARAMS = _o
SHAPES = {"W1": (H, IN_PAD), "W2": (H, H), "W3": (H, H), "W4": (OUT_PAD, H)}


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    out = {}
    for k, (o, n) in OFF.items():
        v = np.asarray(p[o:o + n], dtype=np.float64)
        out[k] = v.reshape(SHAPES[k]) if k in SHAPES else v
    return out


def gelu(x):
    c = 0.7978845608
    return 0.5 * x * (1.0 + np.tanh(c * (x + 0.044715 * x ** 3)))


def gelu_d(x):
    c = 0.7978845608
    t = np.tanh(c * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * c * (1.0 + 3.0 * 0.044715 * x * x)


def loss_and_grad(p, x, t, quantise=True):
    """x (n, 9), t (n, 2) standardised -> (loss, gradient vector)."""
    q = bf16 if quantise else (lambda v: np.asarray(v, dtype=np.float64))
    P = unpack(p)
    W = {k: q(P[k].astype(np.float32)) if quantise else P[k] for k in SHAPES}
    n = x.shape[0]
    X = np.zeros((n, IN_PAD))
    X[:, :IN] = x
    A0 = q(X.astype(np.float32)) if quantise else X
    zs, acts = [], [A0]
    a = A0
    for k, b in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
        z = a @ W[k].T + P[b]
        if quantise:
            z = z.astype(np.float32).astype(np.float64)
        zs.append(z)
        a = q(gelu(z).astype(np.float32)) if quantise else gelu(z)
        acts.append(a)
    y = a @ W["W4"].T + P["b4"]
    e = np.zeros_like(y)
    e[:, :OUT] = y[:, :OUT] - t
    loss = float(np.sum(e * e) / n)
    dz = 2.0 * e / n
    if quantise:
        dz = q(dz.astype(np.float32))
    g = {}
    g["W4"] = dz.T @ acts[3]
    g["b4"] = dz.sum(axis=0)
    da = dz @ W["W4"]
    for li, (k, b) in zip((2, 1, 0), (("W3", "b3"), ("W2", "b2"), ("W1", "b1"))):
        dz = da * gelu_d(zs[li])
        if quantise:
            dz = q(dz.astype(np.float32))
        g[k] = dz.T @ acts[li]
        g[b] = dz.sum(axis=0)
        da = dz @ W[k]
    out = np.zeros(N_PARAMS)
    for k, (o, nn) in OFF.items():
        out[o:o + nn] = np.asarray(g[k]).reshape(-1)
    out[OFF["W4"][0]:OFF["W4"][0] + OFF["W4"][1]].reshape(OUT_PAD, H)[OUT:] = 0.0
    out[OFF["b4"][0] + OUT:OFF["b4"][0] + OUT_PAD] = 0.0
    return loss, out


def forward(p, x, quantise=True):
    q = bf16 if quantise else (lambda v: np.asarray(v, dtype=np.float64))
    P = unpack(p)
    n = x.shape[0]
    X = np.zeros((n, IN_PAD))
    X[:, :IN] = x
    a = q(X.astype(np.float32)) if quantise else X
    for k, b in (("W1", "b1"), ("W2", "b2"), ("W3", "b3")):
        Wk = q(P[k].astype(np.float32)) if quantise else P[k]
        z = a @ Wk.T + P[b]
        a = q(gelu(z).astype(np.float32)) if quantise else gelu(z)
    W4 = q(P["W4"].astype(np.float32)) if quantise else P["W4"]
    return (a @ W4.T + P["b4"])[:, :OUT]
