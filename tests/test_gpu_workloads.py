"""The batched workload at its full size: 256 instances (Kuhn cubes with n
log-uniform in [4, 24], 3.67 M tets in all, seed 5), random plane waves,
eps per 4^3 block of cubes (workloads.batch_instances), advanced together
in captured-graph steps; RK3 and the D3Q4 relaxation scheme.

Checked per instance: the binary32 batch against the binary64 batch
(relative L2 <= 2e-5 after the same steps), and the binary64 batch against
the oracle on the smallest instances (relative L2 <= 1e-12)."""
import numpy as np
import pytest

import gdon3d as G
import oracle as O
from helpers import kuhn_msh, rel_l2
from workloads import batch_instances

STEPS = 5


@pytest.mark.gpu
@pytest.mark.parametrize("kin", [None, "d3q4"])
def test_batched_workload_256_instances(kin):
    inst = batch_instances(256)
    cfl = 5.0 if kin else 0.5
    sims = []
    for n, par, _ in inst:
        sim = G.Simulation(G.Mesh(kuhn=(n, 0.0, 0)), G.Model("plane_wave", kinetic=kin))
        sim.set_time(1.0, 0.0, cfl)
        sims.append(sim)
    pars = np.array([p for _, p, _ in inst])
    eps = np.concatenate([e for _, _, e in inst])
    print(f"256 instances {kin or 'rk3'}: {eps.size} tets", flush=True)
    out = {}
    for prec in ("fp64", "fp32"):
        b = G.Batch(sims, pars, eps=eps, precision=prec)
        b.advance(STEPS)
        b.finish()
        err = b.errors()
        out[prec] = ([s.wn() for s in sims], err, b.times())
        b.free()
        print(f"  {prec}: tiles done, L2 error vs the eps=1 wave median {np.median(err):.3e}",
              flush=True)
    for s in sims:
        s.free()
    w64, e64, t64 = out["fp64"]
    w32, e32, t32 = out["fp32"]
    assert np.array_equal(t64, t32)
    assert np.all(np.isfinite(e64)) and np.all(np.isfinite(e32))
    r = np.array([rel_l2(a, b) for a, b in zip(w32, w64)])
    print(f"  fp32 vs fp64 per instance: max {r.max():.3e}, median {np.median(r):.3e}", flush=True)
    assert r.max() <= 2e-5
    # the smallest instances against the oracle
    small = sorted(range(len(inst)), key=lambda i: inst[i][0])[:3]
    for i in small:
        n, par, e = inst[i]
        om = O.OMesh(kuhn_msh(n, 0.0, 0))
        s = O.OSim(om, 1 if kin else 0, O.PLANE_WAVE, par, kin=O.D3Q4 if kin else 0, eps=e)
        s.set_time(1.0, 0.0, cfl)
        s.solve(STEPS)
        ro = rel_l2(w64[i], s.wn())
        s.free()
        om.free()
        print(f"  instance {i} (n={n}) vs oracle: {ro:.3e}", flush=True)
        assert ro <= 1e-12
