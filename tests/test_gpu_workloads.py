"""The batched workload at its full size: 256 instances (Kuhn cubes with n
log-uniform in [4, 24], 3.67 M tets in all, seed 5), random plane waves,
eps per 4^3 block of cubes (workloads.batch_instances), advanced together
in captured-graph steps; RK3 and the D3Q4 relaxation scheme.

Checked per instance: the binary32 batch against the binary64 batch
(relative L2 <= 2e-5 after the same steps), and the binary64 batch against
the oracle on the smallest instances (relative L2 <= 1e-12); the instances
of one eps block carry the exact wave of speed 1/sqrt(eps), and their L2
errors against it agree between the precisions."""
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
        uni = np.array([n <= 4 for n, _, _ in inst])
        print(f"  {prec}: L2 error of the {uni.sum()} uniform-eps instances against their exact "
              f"wave: median {np.median(err[uni]):.3e}, max {err[uni].max():.3e}", flush=True)
    for s in sims:
        s.free()
    w64, e64, t64 = out["fp64"]
    w32, e32, t32 = out["fp32"]
    assert np.array_equal(t64, t32)
    assert np.all(np.isfinite(e64)) and np.all(np.isfinite(e32))
    r = np.array([rel_l2(a, b) for a, b in zip(w32, w64)])
    print(f"  fp32 vs fp64 per instance: max {r.max():.3e}, median {np.median(r):.3e}", flush=True)
    assert r.max() <= 2e-5
    # the uniform-eps instances' errors against their exact wave: the same
    # in both precisions
    uni = np.array([n <= 4 for n, _, _ in inst])
    assert uni.any() and np.all(np.abs(e32[uni] - e64[uni]) <= 1e-4 * e64[uni] + 1e-7)
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


@pytest.mark.gpu
def test_learned_form_48cubed():
    """The learned-preconditioner workload at its full size: D3Q4 relaxation
    on Kuhn 48^3 x 6 = 663,552 tets with +-20 % jitter (seed 23), eps in
    [1, 4] per element (seed 23), a Gaussian pulse E_z (sigma 0.08), cfl 10
    (20x the explicit CFL). The block network is trained on 10^5 of the
    mesh's items; Richardson block solves with it on 10^4 other items reach
    1e-8 in a few iterations with under 1 % exact fallbacks (the identity
    preconditioner mostly falls back); the sweep on the network's refined
    inverses gives the solution of the stored exact inverses to relative L2
    <= 1e-12 after 10 steps from the same initial state, with the count of
    items kept on the stored inverse reported; the energy stays bounded."""
    from workloads import learned_eps
    G.set_gaussian_pulse((0.5, 0.5, 0.5), 0.08, 1.0)
    mesh = G.Mesh(kuhn=(48, 0.2, 23))
    nelem = mesh.nelem
    sim = G.Simulation(mesh, G.Model("gaussian_pulse", kinetic="d3q4"))
    sim.set_time(1.0, 0.0, 10.0)
    st = G.Stepper(sim, eps=learned_eps(nelem, 23))
    lagged = st.transport_stats()[0]
    perm = np.random.default_rng(23).permutation(4 * nelem)
    x, y = G.learned_data(st, perm[:100000])
    net = G.BlockNet(seed=5)
    net.set_normalization(G.normalization(x, y))
    hist = net.train(x, y, steps=2000, batch=1024, lr=2e-3)
    stats = G.learned_solve_stats(st, net, perm[100000:110000])
    ident = G.learned_solve_stats(st, None, perm[100000:110000])
    print(f"48^3 learned form: {lagged} lagged faces, loss {hist[:5].mean():.3e} -> "
          f"{hist[-20:].mean():.3e}; Richardson mean {stats[0]:.2f} it, max {stats[1]:.0f}, "
          f"fallback {stats[2]:.4f} (identity {ident[2]:.3f})", flush=True)
    assert hist[-20:].mean() < 0.01 * hist[:5].mean()
    assert stats[0] < 10 and stats[2] < 0.01 and stats[3] < 1e-6 and ident[2] > 0.5
    out = {}
    e0 = None
    for mode in ("exact", "learned"):
        ls = G.learned_enable(st, net if mode == "learned" else None)
        G.load_library().dg_stepper_init(st.h)
        if e0 is None:
            e0 = st.energy()
        st.advance(10)
        st.sync()
        out[mode] = (st.download(), st.energy(), ls)
    net.free()
    st.free()
    sim.free()
    r = rel_l2(out["learned"][0], out["exact"][0])
    ls = out["learned"][2]
    print(f"  items on the network / stored inverse {ls[0]} / {ls[1]}, max deviation "
          f"{ls[2]:.2e}, median max|I - D X0| {ls[3]:.2e}; learned vs exact rel L2 {r:.3e}; "
          f"energy {e0:.6e} -> {out['exact'][1]:.6e}", flush=True)
    assert ls[0] > 0.99 * (ls[0] + ls[1])
    assert r <= 1e-12
    # bounded: with eps jumping at every face the kinetic scheme's energy
    # 1/2 (|D|^2 / eps + |B|^2) first grows by some 10 % (the oracle shows the
    # same transient on small meshes), then decays
    assert out["exact"][1] <= 1.25 * e0
