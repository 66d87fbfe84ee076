"""Relaxation on meshes whose upwind graphs have cycles (Kuhn cubes twisted
about the vertical axis, cells positively oriented). The GPU lags the edges
that close cycles and solves exactly in capacitance form (two sweeps per
step around a dense solve for the lagged values, by an inverse computed on
the device at set-up, or, above the dense form's size limit, GMRES with
sweeps as products); the oracle factors
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


@pytest.mark.gpu
@pytest.mark.parametrize("n,alpha,kin,cfl,sweep", [(3, 3.0, "d3q4", 5.0, None),
                                                   (4, 4.0, "d3q4", 5.0, None),
                                                   (3, 3.0, "d3q6", 5.0, None),
                                                   (4, 4.0, "d3q4", 40.0, None),
                                                   (4, 4.0, "d3q4", 5.0, "cluster")])
def test_matrix_free_cycle_solve_matches_exact_lu(n, alpha, kin, cfl, sweep, monkeypatch):
    """Above the dense form's limit on lagged nodes per velocity (lowered to
    0 here with GDN_CAP_MAX_NODES) the lagged values come from GMRES with
    sweeps as its products: more sweeps per step, the same solution as the
    oracle's exact factorization (relative L2 <= 1e-12), also at a large
    time step (CFL 40: values go round the cycles nearly unattenuated) and
    in the cluster sweep."""
    print(f"matrix-free cycles n={n} twist={alpha} {kin} cfl={cfl} {sweep}", flush=True)
    monkeypatch.setenv("GDN_CAP_MAX_NODES", "0")
    if sweep:
        monkeypatch.setenv("GDN_SWEEP", sweep)
    path = twisted_msh(n, alpha)
    g = gpu(path, kin, 0.3, cfl)
    o = oracle(path, kin, 0.3, cfl)
    lagged, last, most = g["stats"]
    r = rel_l2(g["wn"], o)
    print(f"  lagged faces {lagged}, sweeps last step {last}, most {most}, "
          f"rel diff {r:.3e}", flush=True)
    assert lagged > 0 and last > 2
    assert r <= TOL


@pytest.mark.gpu
def test_fp32_matrix_free_cycle_solve(monkeypatch):
    monkeypatch.setenv("GDN_CAP_MAX_NODES", "0")
    path = twisted_msh(4, 4.0)
    a = gpu(path, "d3q4", 0.3, 5.0, "fp32")
    b = gpu(path, "d3q4", 0.3, 5.0, "fp64")
    print(f"  fp32 sweeps last step {a['stats'][1]}, fp64 {b['stats'][1]}, "
          f"rel diff {rel_l2(a['wn'], b['wn']):.3e}", flush=True)
    assert a["stats"][0] == b["stats"][0] > 0
    assert rel_l2(a["wn"], b["wn"]) <= 2e-5


@pytest.mark.gpu
def test_matrix_free_cycle_solve_above_the_dense_limit():
    """A twisted 20^3 Kuhn mesh (48000 elements, strong components of about
    2300 elements per velocity, more lagged nodes in a velocity than the
    dense form takes by default): the lagged values come from GMRES. One
    transport step at CFL 100 with the relaxation switched off (omega = 0)
    satisfies the reference's system M1 x = M2 f of every (velocity, field)
    to a relative residual <= 1e-12, f being the initial state with its
    boundary DoFs at t + dt/2 (oracle assembly left unfactored: dense
    blocks of its strong components would not fit a test)."""
    path = twisted_msh(20, 6.0)
    print("matrix-free cycles above the dense limit: twisted 20^3", flush=True)
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    mesh = G.Mesh(path=path)
    mdl = G.Model("cavity_mode", kinetic="d3q4")
    mdl.s.omega = 0.0
    sim = G.Simulation(mesh, mdl)
    sim.set_time(0.3, 0.0, 100.0)
    st = G.Stepper(sim)
    st.advance(1)
    st.sync()
    stats = st.transport_stats()
    st.finish()
    x = sim.wn()
    st.free()
    sim.free()
    om = O.OMesh(path)
    s = O.OSim(om, 1, O.CAVITY_MODE, cavity_par(), kin=O.D3Q4, factor=False)
    s.set_time(0.3, 0.0, 100.0)
    s.init_solution()
    s.update_boundary(0.5 * s.dt)
    f = s.wn()
    n, nint, nv = s.neq, 10 * om.nelem, 4
    worst = 0.0
    for im in range(6 * nv):
        fi = f[im * n:(im + 1) * n]
        xi = x[im * n:(im + 1) * n].copy()
        xi[nint:] = fi[nint:]  # boundary DoFs: identity rows, the values swept from
        worst = max(worst, s.transport_residual(im % nv, fi, xi))
    s.free()
    om.free()
    lagged, last, most = stats
    print(f"  lagged faces {lagged}, sweeps in the step {last}, largest relative residual "
          f"{worst:.3e}", flush=True)
    assert lagged > 0 and last > 2
    assert worst <= 1e-12


@pytest.mark.gpu
def test_dense_inverse_equals_gmres_on_a_larger_cycle_set(monkeypatch):
    """Twisted 12^3 (601 lagged faces, about 1500 lagged nodes in the
    largest velocity): the dense inverse (Gauss-Jordan on the device) and
    GMRES give the same solution to relative L2 <= 1e-13 after a whole
    solve, with 2 sweeps per step for the dense form."""
    path = twisted_msh(12, 4.0)
    out = {}
    for cap in ("0", "16384"):
        monkeypatch.setenv("GDN_CAP_MAX_NODES", cap)
        out[cap] = gpu(path, "d3q4", 0.1, 5.0)
    r = rel_l2(out["16384"]["wn"], out["0"]["wn"])
    print(f"  dense {out['16384']['stats']}, gmres {out['0']['stats']}, rel diff {r:.3e}",
          flush=True)
    assert out["16384"]["stats"][1] == 2 and out["0"]["stats"][1] > 2
    assert r <= 1e-13
