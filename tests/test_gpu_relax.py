"""GPU parity of the CFL-less relaxation solve (dg_solve_relaxation through
the C ABI: upwind block sweep + node relaxation) against the oracle's
assembled theta-scheme with sparse LU (src/dg/relaxation.c:24-185).

Every oracle comparison names the form of the transport sweep it checks
(GDN_SWEEP = level | item | cluster, GDN_RELAX_FUSED = 0 | 1 | 2: the
relaxation as a pass of its own, fused into the next sweep's reads, or done
by each element's last item) and asserts that the library ran that form;
the benchmark runs whichever of them the set-up timing picks.

Tolerance: relative L2 <= 1e-12 on the full micro state wn (FP64)."""
import os

import numpy as np
import pytest

import gdon3d as G
import oracle as O
from helpers import kuhn_msh, plane_wave_par, cavity_par, rel_l2

TOL = 1e-12
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# (GDN_SWEEP, GDN_RELAX_FUSED): every form the library can run; the cluster
# and level sweeps take a relaxed state (relaxation as a pass of its own)
FORMS = [("item", "1"), ("item", "0"), ("item", "2"), ("cluster", "0"), ("level", "0")]
FORM_IDS = [f"{s}-fused{f}" for s, f in FORMS]

_ORACLE_CACHE = {}


def oracle_run(path, kind, par, kin, tend, cfl):
    """The oracle's solve, once per case and process."""
    key = (path, kind, tuple(np.asarray(par).ravel()), kin, tend, cfl)
    if key not in _ORACLE_CACHE:
        om = O.OMesh(path)
        s = O.OSim(om, 1, kind, par, kin=O.D3Q4 if kin == "d3q4" else O.D3Q6)
        s.set_time(tend, 0.0, cfl)
        s.solve()
        _ORACLE_CACHE[key] = dict(wn=s.wn(), err=s.error_l2(), itermax=s.itermax, t=s.t)
        s.free()
        om.free()
    return _ORACLE_CACHE[key]


def gpu_run(path, imposed, kin, tend, cfl, form, monkeypatch):
    """dg_solve_relaxation's sequence (create, init, advance itermax steps,
    final ghost update and download) with the sweep form forced, and the
    form the library actually ran."""
    sweep, fused = form
    monkeypatch.setenv("GDN_SWEEP", sweep)
    monkeypatch.setenv("GDN_RELAX_FUSED", fused)
    mesh = G.Mesh(path=path)
    sim = G.Simulation(mesh, G.Model(imposed, kinetic=kin))
    sim.set_time(tend, 0.0, cfl)
    st = G.Stepper(sim)
    ran = (G.sweep_info(st)["form"], str(G.relax_fused(st)))
    st.advance(sim.s.itermax)
    st.sync()
    err = st.error_l2()
    st.finish()
    g = dict(wn=sim.wn(), err=err, itermax=sim.s.itermax, t=st.t, form=ran)
    st.free()
    sim.free()
    return g


def run_pair(path, imposed, kind, par, kin, tend, cfl):
    """dg_solve itself (the drop-in entry point, form chosen by the
    library) and the oracle."""
    mesh = G.Mesh(path=path)
    model = G.Model(imposed, kinetic=kin)
    sim = G.Simulation(mesh, model)
    sim.set_time(tend, 0.0, cfl)
    sim.solve()
    g = dict(wn=sim.wn(), err=sim.error_l2(), itermax=sim.s.itermax, t=sim.s.t)
    sim.free()
    return g, oracle_run(path, kind, par, kin, tend, cfl)


def check_form(path, imposed, kind, par, kin, tend, cfl, form, monkeypatch):
    g = gpu_run(path, imposed, kin, tend, cfl, form, monkeypatch)
    o = oracle_run(path, kind, par, kin, tend, cfl)
    r = rel_l2(g["wn"], o["wn"])
    print(f"{os.path.basename(path)} {kin} form {form}: ran {g['form']}, "
          f"relative L2 vs oracle {r:.3e}", flush=True)
    assert g["form"] == form
    assert g["itermax"] == o["itermax"]
    assert r <= TOL
    assert abs(g["err"] - o["err"]) <= 1e-10 * o["err"]


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("n,jitter,tend,cfl", [(3, 0.0, 0.3, 5.0), (4, 0.15, 0.3, 5.0),
                                               (6, 0.0, 0.5, 5.0)])
def test_relax_cavity_d3q4_forms(n, jitter, tend, cfl, form, monkeypatch):
    """Each sweep form against the oracle on the cavity mode: the 3^3 and
    6^3 Kuhn cubes and the 4^3 cube with +-15 % jitter."""
    path = kuhn_msh(n, jitter, 23)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    check_form(path, "cavity_mode", O.CAVITY_MODE, cavity_par(), "d3q4", tend, cfl, form,
               monkeypatch)


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("kin", ["d3q4", "d3q6"])
def test_relax_reference_mesh_forms(kin, form, monkeypatch):
    """Each sweep form against the oracle on the reference's own
    unstructured mesh data/mesh/t4_cube_8.msh (1808 tets)."""
    path = os.path.join(GOLDEN, "t4_cube_8.msh")
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    check_form(path, "cavity_mode", O.CAVITY_MODE, cavity_par(), kin, 0.2, 5.0, form,
               monkeypatch)


@pytest.mark.gpu
@pytest.mark.parametrize("n,jitter,tend,cfl", [(3, 0.0, 0.3, 5.0), (4, 0.15, 0.3, 5.0)])
def test_dg_solve_relaxation_cavity(n, jitter, tend, cfl):
    """dg_solve (the drop-in entry point, form chosen by timing) against
    the oracle."""
    path = kuhn_msh(n, jitter, 23)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    g, o = run_pair(path, "cavity_mode", O.CAVITY_MODE, cavity_par(), "d3q4", tend, cfl)
    assert g["itermax"] == o["itermax"]
    assert rel_l2(g["wn"], o["wn"]) <= TOL
    assert abs(g["err"] - o["err"]) <= 1e-10 * o["err"]


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_relax_plane_wave_d3q4_jittered(form, monkeypatch):
    path = kuhn_msh(4, 0.2, 7)
    k = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2)
    G.set_plane_wave(k, (0, 0, 1))
    check_form(path, "plane_wave", O.PLANE_WAVE, plane_wave_par(k, (0, 0, 1)), "d3q4", 0.3,
               2.0, form, monkeypatch)


@pytest.mark.gpu
@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_relax_d3q6(form, monkeypatch):
    path = kuhn_msh(3, 0.1, 9)
    G.set_plane_wave((2 * np.pi, 0, 0), (0, 1, 0))
    check_form(path, "plane_wave", O.PLANE_WAVE, plane_wave_par(), "d3q6", 0.25, 3.0, form,
               monkeypatch)


@pytest.mark.gpu
@pytest.mark.parametrize("poll", ["sentinel", "full"])
@pytest.mark.parametrize("n,jitter,tmax", [(4, 0.0, 0.6), (6, 0.2, 0.6), (16, 0.1, 0.1)])
def test_persistent_sweep_equals_level_sweep(n, jitter, tmax, poll, monkeypatch):
    """The one-launch dataflow sweep (epoch-tagged hand-off between
    workgroups) computes bit for bit what the level-by-level launches
    compute, with either way of waiting (one word per inflow face, then the
    traces; or the whole traces on every pass). At 16^3 (98,304 items)
    every resident block processes about 20 items, so the in-order item
    pipeline of each block is exercised."""
    path = kuhn_msh(n, jitter, 31)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    monkeypatch.setenv("GDN_SWEEP_POLL", poll)
    out = {}
    for mode in ("level", "persistent"):
        monkeypatch.setenv("GDN_SWEEP", mode)
        mesh = G.Mesh(path=path)
        sim = G.Simulation(mesh, G.Model("cavity_mode", kinetic="d3q4"))
        sim.set_time(tmax, 0.0, 5.0)
        sim.solve()
        out[mode] = sim.wn()
        sim.free()
    assert np.array_equal(out["level"], out["persistent"])


@pytest.mark.gpu
@pytest.mark.parametrize("kin,fused", [("d3q4", "0"), ("d3q4", "1"), ("d3q6", "0")])
def test_item_order_bit_identical(kin, fused, monkeypatch):
    """The sweep's item order (GDN_SWEEP_ORDER: every velocity's upwind
    levels stretched over the deepest velocity's and dealt, or the earliest
    levels) changes only when each item runs: the persistent sweep, the
    level-by-level launches and the cluster sweep each give the same state
    bit for bit in both orders (Kuhn 8^3 cube, where one D3Q4 velocity has
    22 levels and the others 52, and a jittered 6^3 D3Q6 mesh)."""
    path = kuhn_msh(8, 0.0, 0) if kin == "d3q4" else kuhn_msh(6, 0.15, 13)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    monkeypatch.setenv("GDN_RELAX_FUSED", fused)
    for mode in ("item", "level", "cluster"):
        if mode != "item" and fused == "1":
            continue
        monkeypatch.setenv("GDN_SWEEP", mode)
        out = {}
        for order in ("level", "spread"):
            monkeypatch.setenv("GDN_SWEEP_ORDER", order)
            mesh = G.Mesh(path=path)
            sim = G.Simulation(mesh, G.Model("cavity_mode", kinetic=kin))
            sim.set_time(0.2, 0.0, 5.0)
            sim.solve()
            out[order] = sim.wn()
            sim.free()
        print(f"{kin} {mode} fused {fused}: max |spread - level| = "
              f"{np.abs(out['spread'] - out['level']).max():.3e}", flush=True)
        assert np.array_equal(out["level"], out["spread"])


@pytest.mark.gpu
@pytest.mark.parametrize("kin,precision", [("d3q4", "fp64"), ("d3q4", "fp32"), ("d3q6", "fp64")])
def test_fused_relaxation_equals_separate(kin, precision, monkeypatch):
    """The relaxation applied by the next sweep as it reads its input
    (GDN_RELAX_FUSED=1) and as a pass of its own after each sweep
    (GDN_RELAX_FUSED=0) give the same solve: both round the relaxed values
    to the state precision, and the arithmetic differs only in how the
    compiler contracts it (relative L2 <= 1e-13 FP64, <= 1e-5 FP32); the
    fused form also through several steps, the L2 error, the download and a
    second advance."""
    path = kuhn_msh(6, 0.1, 41)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    monkeypatch.setenv("GDN_SWEEP", "item")
    out = {}
    for fused in ("0", "1"):
        monkeypatch.setenv("GDN_RELAX_FUSED", fused)
        mesh = G.Mesh(path=path)
        sim = G.Simulation(mesh, G.Model("cavity_mode", kinetic=kin))
        sim.set_time(0.4, 0.0, 5.0)
        st = G.Stepper(sim, precision=precision)
        assert G.relax_fused(st) == int(fused)
        assert G.sweep_info(st)["form"] == "item"
        half = sim.s.itermax // 2
        st.advance(half)
        err_mid = st.error_l2()
        st.finish()
        mid = sim.wn()
        st.advance(sim.s.itermax - half)
        st.sync()
        err = st.error_l2()
        st.finish()
        out[fused] = (mid, sim.wn(), err_mid, err)
        st.free()
        sim.free()
    tol = 1e-13 if precision == "fp64" else 1e-5
    a, b = out["0"], out["1"]
    print(f"fused vs separate relaxation {kin} {precision}: mid {rel_l2(b[0], a[0]):.3e} "
          f"end {rel_l2(b[1], a[1]):.3e}", flush=True)
    assert rel_l2(b[0], a[0]) <= tol and rel_l2(b[1], a[1]) <= tol
    assert abs(b[2] - a[2]) <= tol * a[2] and abs(b[3] - a[3]) <= tol * a[3]


@pytest.mark.gpu
@pytest.mark.parametrize("n,jitter,tmax,kin,precision", [
    (4, 0.0, 0.6, "d3q4", "fp64"), (6, 0.2, 0.6, "d3q4", "fp64"), (16, 0.1, 0.1, "d3q4", "fp64"),
    (8, 0.3, 0.3, "d3q4", "fp64"), (5, 0.1, 0.4, "d3q6", "fp64"), (6, 0.2, 0.4, "d3q4", "fp32"),
    (16, 0.1, 0.1, "d3q4", "fp32")])
def test_cluster_sweep_equals_level_sweep(n, jitter, tmax, kin, precision, monkeypatch):
    """The cluster sweep (the elements of one grid cell of one velocity
    solved in level order by one wavefront, values inside the cluster passed
    through LDS, values between clusters through the epoch-tagged hand-off)
    computes what the level-by-level launches compute, in both precisions
    and both velocity sets. Its block solve is written directly,
    D^-1 (2 f - sum_f c_f ML_f u_f) - f, where the level and item sweeps
    form D^-1 ML_s before waiting, so the two agree to rounding: relative L2
    <= 1e-13 (FP64), <= 1e-5 (FP32, state rounded to binary32 every step).
    At 16^3 every resident workgroup takes several clusters, so the
    look-ahead of descriptors and headers across clusters is exercised; at
    +-30 % jitter some clusters are split."""
    path = kuhn_msh(n, jitter, 31)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    monkeypatch.setenv("GDN_RELAX_FUSED", "0")
    out = {}
    for mode in ("level", "cluster"):
        monkeypatch.setenv("GDN_SWEEP", mode)
        mesh = G.Mesh(path=path)
        sim = G.Simulation(mesh, G.Model("cavity_mode", kinetic=kin))
        sim.set_time(tmax, 0.0, 5.0)
        st = G.Stepper(sim, precision=precision)
        info = G.sweep_info(st)
        assert info["form"] == mode
        st.advance(sim.s.itermax)
        st.sync()
        st.finish()
        out[mode] = (sim.wn(), info)
        st.free()
        sim.free()
    info = out["cluster"][1]
    print(f"cluster sweep {n}^3 +-{jitter} {kin} {precision}: {info}", flush=True)
    assert info["clusters"] > 0 and info["clustered_items"] > 0
    r = rel_l2(out["cluster"][0], out["level"][0])
    print(f"  cluster vs level: relative L2 {r:.3e}", flush=True)
    assert r <= (1e-13 if precision == "fp64" else 1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("kin,n,jitter", [("d3q4", 4, 0.15), ("d3q4", 6, 0.0), ("d3q6", 3, 0.1)])
def test_cluster_sweep_parity_with_oracle(kin, n, jitter, monkeypatch):
    """The cluster sweep against the oracle's assembled theta-scheme with
    sparse LU: relative L2 <= 1e-12 on the full micro state."""
    monkeypatch.setenv("GDN_SWEEP", "cluster")
    monkeypatch.setenv("GDN_RELAX_FUSED", "0")
    path = kuhn_msh(n, jitter, 23)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    g, o = run_pair(path, "cavity_mode", O.CAVITY_MODE, cavity_par(), kin, 0.3, 5.0)
    r = rel_l2(g["wn"], o["wn"])
    print(f"cluster sweep vs oracle {kin} {n}^3 +-{jitter}: {r:.3e}", flush=True)
    assert r <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("mesh,kin,imposed", [("t4_cube_8", "d3q4", "cavity"),
                                              ("t4_cube_8", "d3q6", "cavity"),
                                              ("t4_cube_8_torus_128", "d3q4", "plane")])
def test_cluster_sweep_unstructured_meshes(mesh, kin, imposed, monkeypatch):
    """The reference's own unstructured meshes (fixtures copied from its
    data/mesh): grid cells cut across the flow there, so most clusters are
    split to single elements and the rest are irregular. The cluster sweep
    (forced) against the oracle: relative L2 <= 1e-12."""
    monkeypatch.setenv("GDN_SWEEP", "cluster")
    monkeypatch.setenv("GDN_RELAX_FUSED", "0")
    path = os.path.join(GOLDEN, mesh + ".msh")
    if imposed == "cavity":
        G.set_cavity_mode(1, 1, 1, (1, -1, 0))
        g, o = run_pair(path, "cavity_mode", O.CAVITY_MODE, cavity_par(), kin, 0.2, 5.0)
    else:
        k = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2)
        G.set_plane_wave(k, (0, 0, 1))
        g, o = run_pair(path, "plane_wave", O.PLANE_WAVE, plane_wave_par(k, (0, 0, 1)),
                        kin, 0.1, 5.0)
    r = rel_l2(g["wn"], o["wn"])
    print(f"cluster sweep vs oracle on {mesh} {kin}: {r:.3e}", flush=True)
    assert r <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2])
@pytest.mark.parametrize("form", FORMS, ids=FORM_IDS)
def test_smallest_meshes_every_sweep_form(n, form, monkeypatch):
    """Edge sizes: one Kuhn cube (6 tets, one cluster per velocity, every
    element next to the boundary) and 2^3 cubes, every form of the sweep
    against the oracle: relative L2 <= 1e-12."""
    path = kuhn_msh(n, 0.0, 0)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    check_form(path, "cavity_mode", O.CAVITY_MODE, cavity_par(), "d3q4", 0.5, 5.0, form,
               monkeypatch)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("form", [("item", "0"), ("cluster", "0"), ("level", "0")],
                         ids=["item", "cluster", "level"])
@pytest.mark.parametrize("kin", ["d3q4", "d3q6"])
def test_relax_projector_boundary_state(kin, form, precision, monkeypatch):
    """The boundary state of the reference's get_fR with use_feq = false
    (src/models/d3q4.c:303-332, src/models/d3q6.c:358-379), selected by
    d3qN_get_fR_projector and computed on the device from the relaxed state
    next to each boundary face, against the oracle's restatement: relative
    L2 <= 1e-12 in binary64 (<= 2e-5 in binary32) on the full micro state
    after whole solves, every sweep form (the relaxation is never fused
    into the sweep here: the ghost update reads the relaxed values)."""
    path = kuhn_msh(4, 0.15, 19)
    k = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2)
    G.set_plane_wave(k, (0, 0, 1))
    sweep, fused = form
    monkeypatch.setenv("GDN_SWEEP", sweep)
    monkeypatch.setenv("GDN_RELAX_FUSED", fused)
    mesh = G.Mesh(path=path)
    sim = G.Simulation(mesh, G.Model("plane_wave", kinetic=kin, boundary_state="projector"))
    sim.set_time(0.3, 0.0, 5.0)
    st = G.Stepper(sim, precision=precision)
    ran = (G.sweep_info(st)["form"], str(G.relax_fused(st)))
    st.advance(sim.s.itermax)
    st.finish()
    wn = sim.wn()
    st.free()
    sim.free()
    om = O.OMesh(path)
    s = O.OSim(om, 1, O.PLANE_WAVE, plane_wave_par(k, (0, 0, 1)),
               kin=O.D3Q4 if kin == "d3q4" else O.D3Q6, fr_projector=True)
    s.set_time(0.3, 0.0, 5.0)
    s.solve()
    ref = s.wn()
    # the f_eq boundary state gives another solution: the projector one is
    # what was run
    s2 = O.OSim(om, 1, O.PLANE_WAVE, plane_wave_par(k, (0, 0, 1)),
                kin=O.D3Q4 if kin == "d3q4" else O.D3Q6)
    s2.set_time(0.3, 0.0, 5.0)
    s2.solve()
    other = s2.wn()
    s.free()
    s2.free()
    om.free()
    r = rel_l2(wn, ref)
    print(f"projector boundary {kin} {form} {precision}: ran {ran}, rel L2 vs oracle {r:.3e}, "
          f"f_eq-boundary oracle differs by {rel_l2(other, ref):.3e}", flush=True)
    assert ran == (sweep, "0")
    assert rel_l2(other, ref) > 1e-6
    assert r <= (TOL if precision == "fp64" else 2e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("fused", ["1", "0"])
@pytest.mark.parametrize("kin,n,jit", [("d3q4", 6, 0.0), ("d3q6", 4, 0.15)])
def test_item_queue_equals_static_order(kin, n, jit, fused, monkeypatch):
    """GDN_SWEEP_QUEUE=dynamic: the item sweep's blocks take their items
    from sharded counters instead of the static order b, b + G, ...; every
    item computes the same values whichever block takes it, so the whole
    solve is bit-identical to the static order (itself checked against the
    oracle above), D3Q4 and D3Q6, relaxation fused or not."""
    path = kuhn_msh(n, jit, 11)
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    out = {}
    for q in ("static", "dynamic"):
        monkeypatch.setenv("GDN_SWEEP_QUEUE", q)
        out[q] = gpu_run(path, "cavity_mode", kin, 0.3, 5.0, ("item", fused), monkeypatch)
    print(f"queue {kin} n={n} fused {fused}: ran {out['dynamic']['form']}", flush=True)
    assert out["static"]["form"] == out["dynamic"]["form"] == ("item", fused)
    assert np.array_equal(out["static"]["wn"], out["dynamic"]["wn"])


def _kuhn_minus_last_tet(n, jitter, seed):
    """A Kuhn n^3 x 6 mesh file without its last tetrahedron: an odd number
    of elements (the remaining cells stay valid; the removed cell's faces
    become boundary faces)."""
    src = kuhn_msh(n, jitter, seed)
    path = src.replace(".msh", "_odd.msh")
    if not os.path.exists(path):
        lines = open(src).read().split("\n")
        i = lines.index("$Elements")
        count = int(lines[i + 1])
        tets = [k for k in range(i + 2, i + 2 + count) if lines[k].split()[1] == "4"]
        del lines[tets[-1]]
        lines[i + 1] = str(count - 1)
        with open(path, "w") as f:
            f.write("\n".join(lines))
    return path


@pytest.mark.gpu
@pytest.mark.parametrize("kin,precision", [("d3q4", "fp64"), ("d3q4", "fp32"), ("d3q6", "fp64"),
                                           ("d3q6", "fp32")])
def test_relaxation_pass_pair_equals_wave(kin, precision, monkeypatch):
    """The relaxation pass with two elements per wave and one 16-byte
    (binary64) or 8-byte (binary32) access per lane and velocity
    (GDN_RELAX_PASS=pair, the default) performs per value the arithmetic of
    the pass with one element per wave (GDN_RELAX_PASS=wave): whole solves
    with the pass after every sweep agree bit for bit, on a mesh with an odd
    number of elements (the last wave holds one) and permittivity per
    element (the flux on D / eps)."""
    path = _kuhn_minus_last_tet(5, 0.1, 7)
    mesh = G.Mesh(path=path)
    ne = mesh.nelem
    mesh.free()
    assert ne % 2 == 1
    eps = np.random.default_rng(5).uniform(1.0, 4.0, ne)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    monkeypatch.setenv("GDN_SWEEP", "item")
    monkeypatch.setenv("GDN_RELAX_FUSED", "0")
    out = {}
    for kind in ("wave", "pair"):
        monkeypatch.setenv("GDN_RELAX_PASS", kind)
        mesh = G.Mesh(path=path)
        sim = G.Simulation(mesh, G.Model("cavity_mode", kinetic=kin))
        sim.set_time(0.3, 0.0, 5.0)
        st = G.Stepper(sim, precision=precision, eps=eps)
        assert G.relax_fused(st) == 0
        st.advance(sim.s.itermax)
        st.sync()
        st.finish()
        out[kind] = sim.wn()
        st.free()
        sim.free()
    print(f"relaxation pass pair vs wave {kin} {precision} ({ne} elements): max |diff| "
          f"{np.abs(out['pair'] - out['wave']).max():.3e}", flush=True)
    assert np.all(np.isfinite(out["pair"]))
    assert np.array_equal(out["pair"], out["wave"])


@pytest.mark.gpu
@pytest.mark.parametrize("kin,eps", [("d3q4", False), ("d3q4", True), ("d3q6", False)])
def test_last_item_relaxation_equals_separate(kin, eps, monkeypatch):
    """The relaxation done inside the sweep by each element's last item
    (GDN_RELAX_FUSED=2: it takes its sibling velocities' transport output
    from their published traces and old values, x = (w - (1 - theta) f) /
    theta, relaxes the element and stores it; no pass) against the pass after
    the sweep (GDN_RELAX_FUSED=0): the recovered output differs from the
    stored one by at most a few units in the last place, so whole solves
    agree to relative L2 <= 1e-13 (binary64) through several steps, the L2
    error, the download and a second advance; jittered mesh, with and
    without permittivity per element."""
    path = kuhn_msh(6, 0.1, 43)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    monkeypatch.setenv("GDN_SWEEP", "item")
    ne = 6 * 6 ** 3
    e = np.random.default_rng(9).uniform(1.0, 4.0, ne) if eps else None
    out = {}
    for form in ("0", "2"):
        monkeypatch.setenv("GDN_RELAX_FUSED", form)
        mesh = G.Mesh(path=path)
        sim = G.Simulation(mesh, G.Model("cavity_mode", kinetic=kin))
        sim.set_time(0.4, 0.0, 5.0)
        st = G.Stepper(sim, eps=e)
        assert G.relax_fused(st) == int(form)
        assert G.sweep_info(st)["form"] == "item"
        half = sim.s.itermax // 2
        st.advance(half)
        err_mid = st.error_l2()
        st.finish()
        mid = sim.wn()
        st.advance(sim.s.itermax - half)
        st.sync()
        err = st.error_l2()
        st.finish()
        out[form] = (mid, sim.wn(), err_mid, err)
        st.free()
        sim.free()
    a, b = out["0"], out["2"]
    print(f"last-item vs pass relaxation {kin} eps={eps}: mid {rel_l2(b[0], a[0]):.3e} "
          f"end {rel_l2(b[1], a[1]):.3e}", flush=True)
    assert rel_l2(b[0], a[0]) <= 1e-13 and rel_l2(b[1], a[1]) <= 1e-13
    assert abs(b[2] - a[2]) <= 1e-13 * a[2] and abs(b[3] - a[3]) <= 1e-13 * a[3]
