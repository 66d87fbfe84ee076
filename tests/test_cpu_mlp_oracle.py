"""The network oracle (oracle/mlp.py) pinned by finite differences of its
own loss (float64 path) and by its sensitivity to a planted mistake: a
gradient computed with a transposed W2 must miss the finite differences."""
import numpy as np

import mlp


def setup(seed=0, n=37):
    rng = np.random.default_rng(seed)
    p = rng.normal(scale=0.3, size=mlp.NPARAM)
    p[0:64 * 16].reshape(64, 16)[:, 7:] = 0.0
    x = rng.normal(size=(n, 7)).astype(np.float32)
    y = rng.normal(size=(n, 100)).astype(np.float32)
    norm = np.concatenate([rng.normal(size=7) * 0.1, 1 + rng.random(7),
                           rng.normal(size=100) * 0.1, 1 + rng.random(100)])
    return p, x, y, norm


def test_gradient_matches_finite_differences():
    p, x, y, norm = setup()
    _, g = mlp.loss_and_grad(p, x, y, norm, bf=False)
    rng = np.random.default_rng(1)
    off = {"W1": 0, "b1": 1024, "W2": 1088, "b2": 5184, "W3": 5248, "b3": 12416}
    picks = [off["W1"] + 16 * r + c for r, c in zip(rng.integers(0, 64, 5), rng.integers(0, 7, 5))]
    picks += list(off["b1"] + rng.integers(0, 64, 3)) + list(off["W2"] + rng.integers(0, 4096, 5))
    picks += list(off["b2"] + rng.integers(0, 64, 3)) + list(off["W3"] + rng.integers(0, 6400, 5))
    picks += list(off["b3"] + rng.integers(0, 100, 3))
    h = 1e-6
    for k in picks:
        e = np.zeros_like(p)
        e[k] = h
        fd = (mlp.loss_and_grad(p + e, x, y, norm, bf=False)[0] -
              mlp.loss_and_grad(p - e, x, y, norm, bf=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])


def test_planted_transpose_is_detected():
    p, x, y, norm = setup(2)
    _, g = mlp.loss_and_grad(p, x, y, norm, bf=False)
    q = p.copy()
    W2 = q[1088:1088 + 4096].reshape(64, 64)
    q[1088:1088 + 4096] = W2.T.reshape(-1)
    _, gq = mlp.loss_and_grad(q, x, y, norm, bf=False)
    rel = np.linalg.norm(gq - g) / np.linalg.norm(g)
    assert rel > 0.1


def test_bf16_rounding():
    a = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 1e-3], np.float32)
    r = mlp.bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12
    assert abs(r[4] - 1e-3) < 4e-6
