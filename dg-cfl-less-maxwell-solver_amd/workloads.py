"""The synthetic workloads of the batched and learned-preconditioner forms,
generated the same way by the benchmark and the tests.

Neither has a counterpart in the reference (one simulation per call there,
eps = mu = 1); the generators use numpy's default_rng with the stated seeds.
"""
import numpy as np


def batch_instances(count=256, seed=5):
    """The batched workload: `count` Kuhn cubes with n log-uniform in
    [4, 24], a plane wave of random direction d on the unit sphere with
    |k| = 2 pi m, m in {1, 2, 3}, E0 a random unit vector orthogonal to k,
    and eps uniform in [1, 4] per 4^3 block of cubes. Returns a list of
    (n, par[16], eps[6 n^3]) with the plane-wave parameter block: for the
    cubes of one block (n <= 4, uniform eps) the exact solution, the wave of
    speed 1/sqrt(eps) with H = sqrt(eps) d x E; for the others the wave of
    speed 1 (initial and boundary data only: with eps jumping between
    blocks there is no closed-form solution)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(np.rint(np.exp(rng.uniform(np.log(4.0), np.log(24.0)))))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        m = int(rng.integers(1, 4))
        e0 = np.cross(d, rng.normal(size=3))
        e0 /= np.linalg.norm(e0)
        k = 2 * np.pi * m * d
        par = np.zeros(16)
        par[0:3], par[3:6], par[6:9], par[9] = k, e0, np.cross(d, e0), np.linalg.norm(k)
        nb = (n + 3) // 4
        table = rng.uniform(1.0, 4.0, size=(nb, nb, nb))
        cube = np.arange(6 * n ** 3) // 6  # 6 tets per cube, cube i + n (j + n k)
        i, j, kk = cube % n, (cube // n) % n, cube // (n * n)
        if nb == 1:
            par[6:9] *= np.sqrt(table[0, 0, 0])
            par[9] /= np.sqrt(table[0, 0, 0])
        out.append((n, par, table[i // 4, j // 4, kk // 4]))
    return out


def learned_eps(nelem, seed=23):
    """The learned form's permittivity: uniform in [1, 4] per element."""
    return np.random.default_rng(seed).uniform(1.0, 4.0, nelem)
