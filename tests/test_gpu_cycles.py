"""Relaxation on meshes whose upwind graphs have cycles (Kuhn cubes twisted
about the vertical axis, cells positively oriented). The GPU lags the edges
that close cycles and solves exactly in capacitance form (two sweeps per
step around a small dense solve for the lagged values); the oracle factors
the assembled M1 exactly (strong components as dense blocks, as KLU's block
triangular form does).

Tolerance: relative L2 <= 1e-12 on the full state (FP64). The FP32 form is
compared with the FP64 GPU form at relative L2 <= 2e-5."""
import numpy as np
import pytest

import gdon3d as G
import oracle as O
from helpers import twisted_msh, cavity_par, rel_l2

TOL = 1e-12


def gpu(path, kin, tend, cfl, precision="fp64"):
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    mesh = G.Mesh(path=path)
    sim = G.Simulation(mesh, G.Model("cavity_mode", kinetic=kin))
    sim.set_time(tend, 0.0, cfl)
    st = G.Stepper(sim, precision=precision)
    st.advance(sim.s.itermax)
    st.sync()
    stats = st.transport_stats()
    st.finish()
    out = dict(wn=sim.wn(), stats=stats, itermax=sim.s.itermax)
    st.free()
    sim.free()
    return out


def oracle(path, kin, tend, cfl):
    om = O.OMesh(path)
    s = O.OSim(om, 1, O.CAVITY_MODE, cavity_par(), kin=O.D3Q4 if kin == "d3q4" else O.D3Q6)
    s.set_time(tend, 0.0, cfl)
    s.solve()
    wn = s.wn()
    s.free()
    om.free()
    return wn


@pytest.mark.gpu
@pytest.mark.parametrize("n,alpha,kin", [(3, 3.0, "d3q4"), (4, 4.0, "d3q4"),
                                         (3, 3.0, "d3q6")])
def test_relaxation_with_upwind_cycles_matches_exact_lu(n, alpha, kin):
    print(f"cycles n={n} twist={alpha} {kin}", flush=True)
    path = twisted_msh(n, alpha)
    g = gpu(path, kin, 0.3, 5.0)
    o = oracle(path, kin, 0.3, 5.0)
    lagged, last, most = g["stats"]
    print(f"  lagged faces {lagged}, sweeps last step {last}, most {most}, "
          f"rel diff {rel_l2(g['wn'], o):.3e}", flush=True)
    assert lagged > 0 and last == 2
    assert rel_l2(g["wn"], o) <= TOL


@pytest.mark.gpu
def test_fp32_relaxation_with_upwind_cycles():
    path = twisted_msh(4, 4.0)
    a = gpu(path, "d3q4", 0.3, 5.0, "fp32")
    b = gpu(path, "d3q4", 0.3, 5.0, "fp64")
    assert a["stats"][0] == b["stats"][0] > 0
    assert rel_l2(a["wn"], b["wn"]) <= 2e-5


@pytest.mark.gpu
def test_cycles_persistent_equals_level(monkeypatch):
    path = twisted_msh(4, 4.0)
    out = {}
    for mode in ("level", "persistent"):
        monkeypatch.setenv("GDN_SWEEP", mode)
        out[mode] = gpu(path, "d3q4", 0.2, 5.0)
    assert out["level"]["stats"] == out["persistent"]["stats"]
    assert np.array_equal(out["level"]["wn"], out["persistent"]["wn"])


@pytest.mark.gpu
def test_cycles_cluster_equals_level(monkeypatch):
    """Over upwind cycles (two sweeps per step, lagged traces given) the
    cluster sweep computes what the level launches compute, to rounding
    (its block solve is written directly; relative L2 <= 1e-13); the
    clusters that lie on a cycle of the cluster graph are solved as single
    elements."""
    path = twisted_msh(4, 4.0)
    out = {}
    for mode in ("level", "cluster"):
        monkeypatch.setenv("GDN_SWEEP", mode)
        out[mode] = gpu(path, "d3q4", 0.2, 5.0)
    assert out["level"]["stats"] == out["cluster"]["stats"]
    assert rel_l2(out["cluster"]["wn"], out["level"]["wn"]) <= 1e-13
