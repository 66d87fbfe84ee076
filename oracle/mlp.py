"""CPU oracle of the learned block preconditioner's network (test
infrastructure only: nothing in the product imports it).

Restates, in numpy, the network of csrc/learned.hip on the same flat
parameter vector: W1[64][16] (7 inputs used), b1[64], W2[64][64], b2[64],
W3[112][64] (100 rows used), b3[112]; inputs and targets normalised as
(x - mu) / sd; tanh hidden layers; loss = mean over items and outputs of the
squared normalised error. With bf16=True every product operand is rounded to
bfloat16 (round to nearest even) as the MFMA path rounds it, and the
products are accumulated in float64 (the GPU accumulates in float32), so the
GPU must agree to ~1e-5 relative; with bf16=False the exact float64
gradient, which the bf16 path approximates to ~1e-2.

This is the textbook two-layer backpropagation, written independently of the
HIP kernels (no reference file: the network has no counterpart in the
reference)."""
import numpy as np

NPARAM = 64 * 16 + 64 + 64 * 64 + 64 + 112 * 64 + 112


def split(p):
    p = np.asarray(p, np.float64)
    o = 0
    out = {}
    for name, shape in (("W1", (64, 16)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)),
                        ("W3", (112, 64)), ("b3", (112,))):
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    assert o == NPARAM
    return out


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def loss_and_grad(p, x, y, norm, bf=True):
    """Loss and gradient (flat, NPARAM) of the mean squared normalised error."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    P = split(p)
    norm = np.asarray(norm, np.float64)
    mux, sdx, muy, sdy = norm[:7], norm[7:14], norm[14:114], norm[114:214]
    n = x.shape[0]
    xn = ((np.asarray(x, np.float32) - mux.astype(np.float32)) / sdx.astype(np.float32)).astype(np.float64)
    yn = ((np.asarray(y, np.float32) - muy.astype(np.float32)) / sdy.astype(np.float32)).astype(np.float64)
    W1, W2, W3 = P["W1"][:, :7], P["W2"], P["W3"][:100]
    h1 = np.tanh(r(xn) @ r(W1).T + P["b1"])
    h2 = np.tanh(r(h1) @ r(W2).T + P["b2"])
    yp = r(h2) @ r(W3).T + P["b3"][:100]
    d = yp - yn
    loss = float(np.mean(d * d))
    dy = 2.0 * d / (n * 100)
    g = {k: np.zeros_like(v) for k, v in P.items()}
    g["W3"][:100] = r(dy).T @ r(h2)
    g["b3"][:100] = dy.sum(0)
    dz2 = (r(dy) @ r(W3)) * (1 - h2 * h2)
    g["W2"] = r(dz2).T @ r(h1)
    g["b2"] = dz2.sum(0)
    dz1 = (r(dz2) @ r(W2)) * (1 - h1 * h1)
    g["W1"][:, :7] = r(dz1).T @ r(xn)
    g["b1"] = dz1.sum(0)
    flat = np.concatenate([g[k].reshape(-1) for k in ("W1", "b1", "W2", "b2", "W3", "b3")])
    return loss, flat


def predict(p, x, norm, bf=True):
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    P = split(p)
    norm = np.asarray(norm, np.float64)
    xn = (np.asarray(x, np.float64) - norm[:7]) / norm[7:14]
    h1 = np.tanh(r(xn) @ r(P["W1"][:, :7]).T + P["b1"])
    h2 = np.tanh(r(h1) @ r(P["W2"]).T + P["b2"])
    yp = r(h2) @ r(P["W3"][:100]).T + P["b3"][:100]
    return yp * norm[114:214] + norm[14:114]
