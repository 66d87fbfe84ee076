"""Batched form (include/dg/batch.h): independent simulations of uneven
size advanced together in one device context (union of the meshes,
captured-graph steps, autotuned tiles).

Each instance of an FP64 batch must equal its own FP64 dg_stepper run bit
for bit (same kernels and time step per element; the other GPU tests pin
the stepper to the oracle at 1e-12). The FP32 batch is compared with the
FP64 runs at relative L2 <= 2e-5 per instance. Permittivity (RK3): a
constant eps = 4 instance against the exact plane wave of speed 1/2."""
import numpy as np
import pytest

import gdon3d as G
from helpers import kuhn_msh, rel_l2


def wave(k, e0, eps=1.0):
    """16-value plane-wave block for permittivity eps: omega = |k|/sqrt(eps),
    H0 = sqrt(eps) k^ x E0."""
    k = np.asarray(k, float)
    e0 = np.asarray(e0, float)
    kn = np.linalg.norm(k)
    par = np.zeros(16)
    par[0:3], par[3:6], par[6:9] = k, e0, np.sqrt(eps) * np.cross(k / kn, e0)
    par[9] = kn / np.sqrt(eps)
    return par


def cavity(lmn, a):
    k = np.pi * np.asarray(lmn, float)
    par = np.zeros(16)
    par[0:3], par[3:6], par[6] = k, a, np.pi * np.sqrt(float(np.dot(lmn, lmn)))
    return par


CASES_RK3 = [((3, 0.1, 1), wave((2 * np.pi, 0, 0), (0, 1, 0))),
             ((4, 0.1, 2), wave((0, 2 * np.pi, 2 * np.pi), (1, 0, 0))),
             ((5, 0.0, 0), wave((2 * np.pi, 2 * np.pi, 0), (0, 0, 1)))]
CASES_KIN = [((3, 0.1, 1), cavity((1, 1, 1), (1, -1, 0))),
             ((4, 0.0, 0), cavity((1, 2, 1), (1, 0, -1))),
             ((2, 0.0, 0), cavity((2, 1, 1), (0, 1, -1)))]


def make_sims(cases, kinetic, cfl):
    sims = []
    for (n, jit, seed), _ in cases:
        imposed = "cavity_mode" if kinetic else "plane_wave"
        sim = G.Simulation(G.Mesh(path=kuhn_msh(n, jit, seed)),
                           G.Model(imposed, kinetic=kinetic))
        sim.set_time(1.0, 0.0, cfl)
        sims.append(sim)
    return sims


def single(case, kinetic, cfl, steps):
    (n, jit, seed), par = case
    if kinetic:
        lmn = np.rint(par[0:3] / np.pi).astype(int)
        G.set_cavity_mode(int(lmn[0]), int(lmn[1]), int(lmn[2]), par[3:6])
    else:
        G.set_plane_wave(par[0:3], par[3:6])
    imposed = "cavity_mode" if kinetic else "plane_wave"
    sim = G.Simulation(G.Mesh(path=kuhn_msh(n, jit, seed)), G.Model(imposed, kinetic=kinetic))
    sim.set_time(1.0, 0.0, cfl)
    st = G.Stepper(sim)
    st.advance(steps)
    st.finish()
    wn = sim.wn()
    st.free()
    sim.free()
    return wn


def run_batch(cases, kinetic, cfl, steps, precision, split=None):
    sims = make_sims(cases, kinetic, cfl)
    b = G.Batch(sims, np.array([p for _, p in cases]), precision=precision)
    for k in (split or [steps]):
        b.advance(k)
    b.finish()
    out = [s.wn() for s in sims]
    tiles = b.tiles()
    b.free()
    for s in sims:
        s.free()
    return out, tiles


@pytest.mark.gpu
@pytest.mark.parametrize("kinetic,cases,cfl,fused,rk3,sweep",
                         [(None, CASES_RK3, 0.5, "0", "pair", None),
                          (None, CASES_RK3, 0.5, "0", "tile", None),
                          (None, CASES_RK3, 0.5, "0", "soa", None),
                          ("d3q4", CASES_KIN, 5.0, "0", None, "item"),
                          ("d3q4", CASES_KIN, 5.0, "0", None, "cluster"),
                          ("d3q4", CASES_KIN, 5.0, "1", None, "item")])
def test_fp64_batch_equals_each_instance_alone(kinetic, cases, cfl, fused, rk3, sweep,
                                               monkeypatch):
    """Bit for bit, with the RK3 stage kernel (GDN_RK3: the three kernels
    agree to rounding only), the form of the sweep (GDN_SWEEP: item and
    cluster sweeps agree to rounding only) and the relaxation form
    (separate, or fused into the sweep: GDN_RELAX_FUSED) fixed for the batch
    and the single instances."""
    monkeypatch.setenv("GDN_RELAX_FUSED", fused)
    if rk3:
        monkeypatch.setenv("GDN_RK3", rk3)
    if sweep:
        monkeypatch.setenv("GDN_SWEEP", sweep)
    print(f"batch {kinetic or 'rk3'} fused {fused} rk3 {rk3}", flush=True)
    steps = 7
    out, tiles = run_batch(cases, kinetic, cfl, steps, "fp64")
    print(f"  tiles {tiles}", flush=True)
    for case, wn in zip(cases, out):
        ref = single(case, kinetic, cfl, steps)
        assert np.array_equal(wn, ref), rel_l2(wn, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("kinetic,cases,cfl,fused,sweep",
                         [(None, CASES_RK3, 0.5, "0", None), ("d3q4", CASES_KIN, 5.0, "0", "item"),
                          ("d3q4", CASES_KIN, 5.0, "0", "cluster"),
                          ("d3q4", CASES_KIN, 5.0, "1", "item")])
def test_batch_graph_replay_split_steps(kinetic, cases, cfl, fused, sweep, monkeypatch):
    """7 steps in one call or as 3 + 1 + 2 + 1 (graph pairs of both parities,
    direct odd steps; with the fused relaxation the first step of the batch
    is taken directly so that every captured sweep relaxes its input) give
    the same states bit for bit (the sweep form fixed: GDN_SWEEP)."""
    monkeypatch.setenv("GDN_RELAX_FUSED", fused)
    if sweep:
        monkeypatch.setenv("GDN_SWEEP", sweep)
    a, _ = run_batch(cases, kinetic, cfl, 7, "fp64")
    b, _ = run_batch(cases, kinetic, cfl, 7, "fp64", split=[3, 1, 2, 1])
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("kinetic,cases,cfl", [(None, CASES_RK3, 0.5), ("d3q4", CASES_KIN, 5.0)])
def test_fp32_batch_close_to_fp64_instances(kinetic, cases, cfl):
    steps = 20
    out, _ = run_batch(cases, kinetic, cfl, steps, "fp32")
    for case, wn in zip(cases, out):
        ref = single(case, kinetic, cfl, steps)
        assert rel_l2(wn, ref) <= 2e-5


@pytest.mark.gpu
def test_rk3_permittivity_plane_wave():
    """eps = 4 everywhere: the plane wave of speed 1/2 (omega = |k|/2,
    H0 = 2 k^ x E0) is exact; the error at n = 8 after t ~ 0.5 is of the
    size of the eps = 1 error on the same mesh."""
    errs = {}
    for eps in (1.0, 4.0):
        mesh = G.Mesh(path=kuhn_msh(8, 0.0, 0))
        ne = mesh.nelem
        sim = G.Simulation(mesh, G.Model("plane_wave"))
        sim.set_time(1.0, 0.0, 0.5)
        par = wave((2 * np.pi, 0, 0), (0, 1, 0), eps)
        b = G.Batch([sim], par[None, :], eps=np.full(ne, eps), precision="fp64")
        b.advance(116)
        errs[eps] = float(b.errors()[0])
        b.free()
        sim.free()
    print(f"  errors eps=1 {errs[1.0]:.3e} eps=4 {errs[4.0]:.3e}", flush=True)
    assert errs[4.0] < 5e-3 and errs[1.0] < 5e-3
    assert errs[4.0] < 3 * errs[1.0]


@pytest.mark.gpu
@pytest.mark.parametrize("sweep", ["item", "cluster"])
def test_batch_with_one_cube_instance(sweep, monkeypatch):
    """Ragged batch: an instance of a single Kuhn cube (6 tets) between two
    larger ones; each instance equals its own stepper run bit for bit."""
    monkeypatch.setenv("GDN_RELAX_FUSED", "0")
    monkeypatch.setenv("GDN_SWEEP", sweep)
    cases = [((3, 0.1, 1), cavity((1, 1, 1), (1, -1, 0))),
             ((1, 0.0, 0), cavity((1, 1, 2), (1, 0, -1))),
             ((2, 0.0, 0), cavity((2, 1, 1), (0, 1, -1)))]
    out, _ = run_batch(cases, "d3q4", 5.0, 5, "fp64")
    for case, wn in zip(cases, out):
        ref = single(case, "d3q4", 5.0, 5)
        assert np.array_equal(wn, ref), rel_l2(wn, ref)
