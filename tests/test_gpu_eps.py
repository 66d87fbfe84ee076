"""Permittivity (eps per element, mu = 1; no counterpart in the reference,
whose flux hard-wires eps = mu = 1): GPU against the oracle's restatement
(oracle/oracle.c: RK3 with 1/eps on the E rows of M^-1; kinetic scheme on
the conserved (D, B) with the equilibrium flux on (D / eps, B); both pinned
in tests/test_cpu_oracle.py by the uniform-eps exact wave, the energy
identity and the equilibrium moments). Relative L2 <= 1e-12 on the full
state (binary64), per instance for batches with eps per 4^3 block of cubes
(the batched workload's layout); the Gaussian pulse of the learned form."""
import numpy as np
import pytest

import gdon3d as G
import oracle as O
from helpers import block_eps, gaussian_par, kuhn_msh, plane_wave_eps_par, rel_l2

TOL = 1e-12
K = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2)


def oracle_solve(path, scheme, kind, par, kin, eps, tend, cfl, nsteps=-1):
    om = O.OMesh(path)
    s = O.OSim(om, scheme, kind, par, kin={None: 0, "d3q4": O.D3Q4, "d3q6": O.D3Q6}[kin],
               eps=eps)
    s.set_time(tend, 0.0, cfl)
    s.solve(nsteps)
    out = dict(wn=s.wn(), err=s.error_l2(), energy=s.energy(), itermax=s.itermax)
    s.free()
    om.free()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("form", [("item", "1"), ("item", "0"), ("cluster", "0")],
                         ids=["item-fused1", "item-fused0", "cluster"])
@pytest.mark.parametrize("kin", ["d3q4", "d3q6"])
def test_kinetic_eps_per_element_vs_oracle(kin, form, monkeypatch):
    """eps uniform in [1, 4] per element (seed 11), jittered 4^3 cubes,
    oblique plane wave data, each sweep form forced: state, L2 error (E =
    D / eps against the exact E) and energy 1/2 (|D|^2 / eps + |B|^2)."""
    path = kuhn_msh(4, 0.15, 5)
    ne = 6 * 4 ** 3
    eps = np.random.default_rng(11).uniform(1.0, 4.0, ne)
    G.set_plane_wave(K, (0, 0, 1))
    monkeypatch.setenv("GDN_SWEEP", form[0])
    monkeypatch.setenv("GDN_RELAX_FUSED", form[1])
    sim = G.Simulation(G.Mesh(path=path), G.Model("plane_wave", kinetic=kin))
    sim.set_time(0.3, 0.0, 5.0)
    st = G.Stepper(sim, eps=eps)
    ran = (G.sweep_info(st)["form"], str(G.relax_fused(st)))
    st.advance(sim.s.itermax)
    st.sync()
    err, en = st.error_l2(), st.energy()
    st.finish()
    wn = sim.wn()
    st.free()
    sim.free()
    o = oracle_solve(path, 1, O.PLANE_WAVE, plane_wave_eps_par(K, (0, 0, 1), 1.0), kin, eps,
                     0.3, 5.0)
    r = rel_l2(wn, o["wn"])
    print(f"eps per element {kin} {form}: ran {ran}, rel L2 {r:.3e}, err {err:.6e} / "
          f"{o['err']:.6e}, energy {en:.12e} / {o['energy']:.12e}", flush=True)
    assert ran == form
    assert r <= TOL
    assert abs(err - o["err"]) <= 1e-10 * o["err"]
    assert abs(en - o["energy"]) <= 1e-12 * o["energy"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rk3_eps_per_block_vs_oracle(precision):
    """RK3 with eps per 2^3 block of cubes (seed 3) on jittered 4^3 cubes."""
    path = kuhn_msh(4, 0.1, 6)
    eps = block_eps(path, 2, 3)
    G.set_plane_wave(K, (0, 0, 1))
    sim = G.Simulation(G.Mesh(path=path), G.Model("plane_wave"))
    sim.set_time(0.2, 0.0, 0.5)
    st = G.Stepper(sim, precision=precision, eps=eps)
    st.advance(sim.s.itermax)
    st.sync()
    en = st.energy()
    st.finish()
    wn = sim.wn()
    st.free()
    sim.free()
    o = oracle_solve(path, 0, O.PLANE_WAVE, plane_wave_eps_par(K, (0, 0, 1), 1.0), None, eps,
                     0.2, 0.5)
    r = rel_l2(wn, o["wn"])
    print(f"RK3 eps per block {precision}: rel L2 {r:.3e}, energy {en:.12e} / "
          f"{o['energy']:.12e}", flush=True)
    tol = TOL if precision == "fp64" else 2e-5
    assert r <= tol
    assert abs(en - o["energy"]) <= tol * o["energy"]


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", ["rk3", "d3q4"])
def test_uniform_eps_exact_wave_on_gpu(scheme):
    """eps = 4 everywhere with the exact wave of speed 1/2
    (maxwell_plane_wave_set_eps): the GPU's L2 error equals the oracle's and
    is small (the oracle's convergence to this wave is pinned on the CPU)."""
    path = kuhn_msh(5, 0.0, 0)
    k = np.array([2 * np.pi, 0.0, 0.0])
    G.set_plane_wave_eps(k, (0, 1, 0), 4.0)
    kin = None if scheme == "rk3" else scheme
    cfl = 0.5 if kin is None else 5.0
    sim = G.Simulation(G.Mesh(path=path), G.Model("plane_wave", kinetic=kin))
    sim.set_time(0.3, 0.0, cfl)
    st = G.Stepper(sim, eps=np.full(6 * 5 ** 3, 4.0))
    st.advance(sim.s.itermax)
    st.sync()
    err = st.error_l2()
    st.free()
    sim.free()
    G.set_plane_wave((2 * np.pi, 0, 0), (0, 1, 0))
    o = oracle_solve(path, 0 if kin is None else 1, O.PLANE_WAVE,
                     plane_wave_eps_par(k, (0, 1, 0), 4.0), kin, np.full(6 * 5 ** 3, 4.0),
                     0.3, cfl)
    print(f"uniform eps=4 {scheme}: GPU L2 error {err:.6e}, oracle {o['err']:.6e}", flush=True)
    assert abs(err - o["err"]) <= 1e-10 * o["err"]
    assert err < (3e-2 if kin else 1e-2)


@pytest.mark.gpu
@pytest.mark.parametrize("kin", ["d3q4", None])
def test_gaussian_pulse_with_eps_vs_oracle(kin):
    """The learned form's initial state: a Gaussian pulse E_z (sigma 0.08,
    centred) with eps per element, jittered 4^3 cubes."""
    path = kuhn_msh(4, 0.2, 23)
    eps = np.random.default_rng(23).uniform(1.0, 4.0, 6 * 4 ** 3)
    G.set_gaussian_pulse((0.5, 0.5, 0.5), 0.08, 1.0)
    sim = G.Simulation(G.Mesh(path=path), G.Model("gaussian_pulse", kinetic=kin))
    cfl = 10.0 if kin else 0.5
    sim.set_time(0.2, 0.0, cfl)
    st = G.Stepper(sim, eps=eps)
    e0 = st.energy()
    st.advance(sim.s.itermax)
    st.finish()
    wn = sim.wn()
    st.free()
    sim.free()
    o = oracle_solve(path, 1 if kin else 0, O.GAUSSIAN_PULSE, gaussian_par(), kin, eps, 0.2,
                     cfl)
    r = rel_l2(wn, o["wn"])
    print(f"Gaussian pulse eps {kin}: rel L2 {r:.3e}, energy at t=0 {e0:.6e}", flush=True)
    assert r <= TOL


def batch_cases(kin):
    """three uneven instances, eps per 4^3 block of cubes as the batched
    workload draws it (numpy default_rng)"""
    rng = np.random.default_rng(5)
    out = []
    for n, jit in ((3, 0.1), (5, 0.0), (8, 0.05)):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        e0 = np.cross(d, rng.normal(size=3))
        e0 /= np.linalg.norm(e0)
        m = int(rng.integers(1, 4))
        nb = (n + 3) // 4
        table = rng.uniform(1.0, 4.0, size=(nb, nb, nb))
        cube = np.arange(6 * n ** 3) // 6
        i, j, k = cube % n, (cube // n) % n, cube // (n * n)
        out.append((n, jit, 2 * np.pi * m * d, e0, table[i // 4, j // 4, k // 4]))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("kin", [None, "d3q4"])
def test_batch_eps_per_block_vs_oracle(kin):
    """A batch (captured-graph steps) of three uneven instances with eps
    per 4^3 block: every instance against its own oracle solve."""
    cases = batch_cases(kin)
    cfl = 5.0 if kin else 0.5
    steps = 6
    sims, pars = [], []
    for n, jit, k, e0, _ in cases:
        G.set_plane_wave(k, e0)
        sim = G.Simulation(G.Mesh(path=kuhn_msh(n, jit, 9)), G.Model("plane_wave", kinetic=kin))
        sim.set_time(1.0, 0.0, cfl)
        sims.append(sim)
        pars.append(plane_wave_eps_par(k, e0, 1.0))
    b = G.Batch(sims, np.array(pars), eps=np.concatenate([c[4] for c in cases]),
                precision="fp64")
    b.advance(steps)
    b.finish()
    out = [s.wn() for s in sims]
    b.free()
    for s in sims:
        s.free()
    for (n, jit, k, e0, eps), par, wn in zip(cases, pars, out):
        o = oracle_solve(kuhn_msh(n, jit, 9), 1 if kin else 0, O.PLANE_WAVE, par, kin, eps,
                         1.0, cfl, nsteps=steps)
        r = rel_l2(wn, o["wn"])
        print(f"batch eps {kin} instance n={n}: rel L2 vs oracle {r:.3e}", flush=True)
        assert r <= TOL
