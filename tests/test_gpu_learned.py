"""Learned block preconditioner of the transport solve (csrc/learned.hip).

* Gradients of the GPU network (bf16 MFMA products, fp32 accumulation) are
  checked against the CPU oracle (oracle/mlp.py): with the oracle rounding
  the same operands to bf16, relative L2 <= 2e-3 per parameter block and the
  same loss to 1e-4; against the exact float64 gradient, relative L2 <= 5e-2.
* Training on the block parameters and exact inverses of one mesh lowers the
  loss by 10x or more.
* On another mesh, Richardson block solves preconditioned by the trained
  network reach |r - D u| <= 1e-8 |r| in about 7 iterations on average with
  at most 1 % fallbacks, where the identity preconditioner (always) and the
  untrained network (mostly) fall back to the exact inverse; every result is
  within 1e-6 (relative L2) of the exact solution D^-1 r. Right-preconditioned
  GMRES converges on every solve within 10 products with any of the three
  (measured: identity 6.7 on average, trained 6.9, untrained 10), so the
  network pays only for Richardson.
"""
import numpy as np
import pytest

import gdon3d as G
import mlp as M
from helpers import kuhn_msh


def stepper(n, jit, seed, cfl=10.0):
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    sim = G.Simulation(G.Mesh(path=kuhn_msh(n, jit, seed)),
                       G.Model("cavity_mode", kinetic="d3q4"))
    sim.set_time(1.0, 0.0, cfl)
    return sim, G.Stepper(sim)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


BLOCKS = {"W1": (0, 1024), "b1": (1024, 1088), "W2": (1088, 5184), "b2": (5184, 5248),
          "W3": (5248, 12416), "b3": (12416, 12528)}


@pytest.mark.gpu
def test_gradient_matches_cpu_oracle():
    rng = np.random.default_rng(0)
    net = G.BlockNet(seed=3)
    n = 100  # not a multiple of 16: padded items
    x = rng.normal(size=(n, 7)).astype(np.float32)
    y = rng.normal(size=(n, 100)).astype(np.float32)
    norm = np.concatenate([rng.normal(size=7) * 0.1, 1 + rng.random(7),
                           rng.normal(size=100) * 0.1, 1 + rng.random(100)]).astype(np.float32)
    net.set_normalization(norm)
    loss, g = net.grad(x, y)
    p = net.params()
    lb, gb = M.loss_and_grad(p, x, y, norm, bf=True)
    lx, gx = M.loss_and_grad(p, x, y, norm, bf=False)
    net.free()
    for name, (a, b) in BLOCKS.items():
        print(f"  {name}: rel vs bf16 oracle {rel(g[a:b], gb[a:b]):.2e}, "
              f"vs exact {rel(g[a:b], gx[a:b]):.2e}", flush=True)
        assert rel(g[a:b], gb[a:b]) <= 2e-3
    assert abs(loss - lb) <= 1e-4 * lb
    assert rel(g, gx) <= 5e-2


@pytest.mark.gpu
def test_training_and_what_the_network_gains():
    sim, st = stepper(12, 0.2, 23)
    nitems = 4 * 6 * 12 ** 3
    x, y = G.learned_data(st, np.arange(nitems))
    net = G.BlockNet(seed=5)
    norm = G.normalization(x, y)
    net.set_normalization(norm)
    hist = net.train(x, y, steps=800, batch=1024, lr=2e-3)
    print(f"  loss {hist[:5].mean():.3e} -> {hist[-20:].mean():.3e}", flush=True)
    assert hist[-20:].mean() < 0.1 * hist[:5].mean()
    st.free()
    sim.free()

    sim2, st2 = stepper(10, 0.15, 7)
    idx = np.random.default_rng(1).choice(4 * 6 * 10 ** 3, 4000, replace=False)
    trained = G.learned_solve_stats(st2, net, idx)
    ident = G.learned_solve_stats(st2, None, idx)
    fresh = G.BlockNet(seed=9)
    fresh.set_normalization(norm)
    untrained = G.learned_solve_stats(st2, fresh, idx)
    for name, s in (("trained", trained), ("identity", ident), ("untrained", untrained)):
        print(f"  {name}: mean iterations {s[0]:.2f}, max {s[1]:.0f}, fallback rate "
              f"{s[2]:.3f}, max rel L2 to exact {s[3]:.2e}", flush=True)
    assert trained[2] <= 0.01 and trained[0] <= 12
    assert trained[3] <= 1e-6
    assert ident[2] > 0.5 and untrained[2] > trained[2] + 0.2
    # right-preconditioned GMRES on the same solves: every one converges
    # within 10 products whatever the preconditioner (a 10x10 system; the
    # identity needs about 7 here, the trained network about as many), so
    # the network's gain is Richardson's, not GMRES's
    g = {name: G.learned_solve_stats(st2, p, idx, method="gmres")
         for name, p in (("trained", net), ("identity", None), ("untrained", fresh))}
    for name, st_ in g.items():
        print(f"  GMRES {name}: mean products {st_[0]:.2f}, max {st_[1]:.0f}, fallback rate "
              f"{st_[2]:.3f}, max rel L2 to exact {st_[3]:.2e}", flush=True)
    assert all(st_[2] == 0.0 and st_[3] <= 1e-6 and st_[1] <= 10 for st_ in g.values())
    fresh.free()
    net.free()
    st2.free()
    sim2.free()


def cavity_oracle(path, tend, cfl):
    import oracle as O
    from helpers import cavity_par
    om = O.OMesh(path)
    s = O.OSim(om, 1, O.CAVITY_MODE, cavity_par(), kin=O.D3Q4)
    s.set_time(tend, 0.0, cfl)
    s.solve()
    w = s.wn()
    s.free()
    om.free()
    return w


def trained_net(n=6, jit=0.2, seed=23, steps=800):
    sim, st = stepper(n, jit, seed)
    x, y = G.learned_data(st, np.arange(4 * 6 * n ** 3))
    net = G.BlockNet(seed=5)
    net.set_normalization(G.normalization(x, y))
    net.train(x, y, steps=steps, batch=1024, lr=2e-3)
    st.free()
    sim.free()
    return net


@pytest.mark.gpu
@pytest.mark.parametrize("form", [("item", "1"), ("item", "0"), ("level", "0")],
                         ids=["item-fused1", "item-fused0", "level"])
def test_learned_inverses_in_the_sweep_vs_oracle(form, monkeypatch):
    """The sweep on the network's refined block inverses (trained on one
    mesh, used on another) against the oracle's sparse LU: relative L2 <=
    1e-12 after a whole solve, the count of items kept on the stored exact
    inverse reported (measured: 2170 of 3000 on the network, median
    max|I - D X0| = 0.12); and the same solve with every item on the stored
    exact inverse."""
    from helpers import rel_l2
    net = trained_net()
    path = kuhn_msh(5, 0.15, 3)
    ref = cavity_oracle(path, 0.3, 10.0)
    monkeypatch.setenv("GDN_SWEEP", form[0])
    monkeypatch.setenv("GDN_RELAX_FUSED", form[1])
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    out = {}
    for mode in ("learned", "exact"):
        sim = G.Simulation(G.Mesh(path=path), G.Model("cavity_mode", kinetic="d3q4"))
        sim.set_time(0.3, 0.0, 10.0)
        st = G.Stepper(sim)
        stats = G.learned_enable(st, net if mode == "learned" else None)
        st.lib.dg_stepper_init(st.h)
        ran = (G.sweep_info(st)["form"], str(G.relax_fused(st)))
        st.advance(sim.s.itermax)
        st.finish()
        out[mode] = sim.wn()
        st.free()
        sim.free()
        r = rel_l2(out[mode], ref)
        print(f"  {mode} {form}: ran {ran}, items on the network / stored inverse / max dev "
              f"{stats}, rel L2 vs oracle {r:.3e}", flush=True)
        assert ran == form
        assert r <= 1e-12
        if mode == "learned":
            # most items on the network's inverse, refined to rounding
            assert stats[0] > 0.5 * (stats[0] + stats[1]) and stats[2] < 1e-13
    net.free()


@pytest.mark.gpu
def test_untrained_network_falls_back_to_exact_inverses():
    """An untrained network's predictions do not converge under the
    refinement: those items keep the stored exact inverse (the judge), and
    the solve is still the oracle's."""
    from helpers import rel_l2
    net = G.BlockNet(seed=9)
    sim0, st0 = stepper(4, 0.1, 1)
    x, y = G.learned_data(st0, np.arange(4 * 6 * 4 ** 3))
    net.set_normalization(G.normalization(x, y))
    st0.free()
    sim0.free()
    path = kuhn_msh(4, 0.1, 2)
    ref = cavity_oracle(path, 0.3, 10.0)
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    sim = G.Simulation(G.Mesh(path=path), G.Model("cavity_mode", kinetic="d3q4"))
    sim.set_time(0.3, 0.0, 10.0)
    st = G.Stepper(sim)
    stats = G.learned_enable(st, net)
    st.lib.dg_stepper_init(st.h)
    st.advance(sim.s.itermax)
    st.finish()
    r = rel_l2(sim.wn(), ref)
    print(f"  untrained: items on the network / stored inverse {stats[:2]}, rel L2 {r:.3e}",
          flush=True)
    st.free()
    sim.free()
    net.free()
    assert stats[1] > 0.5 * (stats[0] + stats[1])
    assert r <= 1e-12
