"""Pins of the oracle's assembled operators against what the mathematics
fixes, so that a dropped term, a wrong sign or index, or a transposed
operand fails one of them:

* the RK3 operator dw = M^-1 F w is exact for quadratic fields (the DG
  residual of a global polynomial reduces to sum_k A(e_k) d_k w);
* the upwind operator dissipates energy: w^T M dw >= 0 with zero ghosts;
* LSRK3 converges at order 3 in time (Richardson ratios);
* the theta-scheme transport M1^-1 M2 is exact for fields linear in space
  and time, its LU residual is at round-off, and its sparsity is
  block-triangular with 10x10 element blocks;
* relaxation keeps the macroscopic moments;
* the L2 error of an offset exact interpolant is the offset.
"""
import ctypes as C

import numpy as np
import pytest

import gdon3d as G
import oracle as O
from helpers import kuhn_msh, plane_wave_par, cavity_par

rng = np.random.default_rng(7)


def A(n):
    N = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
    M = np.zeros((6, 6))
    M[:3, 3:] = -N
    M[3:, :3] = N
    return M


def dof_coords(path):
    """coordinates of every DoF (element nodes then ghost nodes)"""
    om = O.OMesh(path)
    a = om.arrays()
    x = a["node"].reshape(-1, 3)[a["elem2qnode"]]
    return om, x


@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_rk3_operator_exact_for_quadratic_fields(jitter):
    path = kuhn_msh(3, jitter, 11)
    om, X = dof_coords(path)
    s = O.OSim(om, 0, O.PLANE_WAVE, plane_wave_par())
    c0 = rng.normal(size=6)
    B = rng.normal(size=(6, 3))
    Q = rng.normal(size=(6, 3, 3))
    Q = Q + Q.transpose(0, 2, 1)
    w = c0 + X @ B.T + 0.5 * np.einsum("nk,qkl,nl->nq", X, Q, X)
    grad = B[None, :, :] + np.einsum("qkl,nl->nqk", Q, X)  # [node, q, k]
    expect = sum(np.einsum("pq,nq->np", A(np.eye(3)[k]), grad[:, :, k]) for k in range(3))
    dw = s.rk3_operator(w.ravel()).reshape(-1, 6)
    ne = om.nelem
    err = np.abs(dw[: 10 * ne] - expect[: 10 * ne]).max() / np.abs(expect).max()
    assert err < 1e-7, err  # O(1) for a wrong term; face mass rounding ~1e-8
    assert np.all(dw[10 * ne:] == 0)  # ghost rows of M^-1 F vanish
    s.free()
    om.free()


def element_mass(path):
    om = O.OMesh(path)
    t = O.tables()
    lib = G.load_library()
    m = G.Mesh(path=path)
    det = np.zeros(1)
    dets = []
    for e in range(m.nelem):
        lib.mesh_get_elem_det(C.byref(m.s), e, G.dptr(det))
        dets.append(det[0])
    m.free()
    return om, t["mass"], np.array(dets)


def test_rk3_operator_dissipates_energy():
    path = kuhn_msh(3, 0.15, 4)
    om, M, det = element_mass(path)
    s = O.OSim(om, 0, O.PLANE_WAVE, plane_wave_par())
    ne = om.nelem
    for _ in range(5):
        w = np.zeros((s.neq, 6))
        w[: 10 * ne] = rng.normal(size=(10 * ne, 6))
        dw = s.rk3_operator(w.ravel()).reshape(-1, 6)
        we, de = w[: 10 * ne].reshape(ne, 10, 6), dw[: 10 * ne].reshape(ne, 10, 6)
        energy_rate = np.einsum("eiq,ij,e,ejq->", we, M, det, de)
        scale = np.einsum("eiq,ij,e,ejq->", we, M, det, we)
        assert energy_rate >= -1e-12 * scale
        assert energy_rate > 1e-6 * scale  # upwinding does dissipate
    s.free()
    om.free()


def test_rk3_third_order_in_time():
    path = kuhn_msh(2, 0.0, 0)
    om = O.OMesh(path)
    sols = []
    for dt in (2.0 ** -5, 2.0 ** -6, 2.0 ** -7):  # t_end = 1/4 exactly
        s = O.OSim(om, 0, O.PLANE_WAVE, plane_wave_par())
        s.set_time(0.25, dt, -1)
        s.solve()
        # interior DoFs only: like the reference, the ghosts keep the data of
        # the last stage time, t_end - dt/4
        sols.append(s.wn()[: 60 * om.nelem])
        s.free()
    r = np.linalg.norm(sols[0] - sols[1]) / np.linalg.norm(sols[1] - sols[2])
    assert 2.8 < np.log2(r) < 3.2, np.log2(r)
    om.free()


@pytest.mark.parametrize("kin,nv", [(O.D3Q4, 4), (O.D3Q6, 6)])
def test_transport_exact_for_linear_fields(kin, nv):
    """f(x, t) = a + b.(x - v t) solves f_t + v.grad f = 0; Crank-Nicolson
    with ghosts at t + dt/2 and P2 in space reproduce it."""
    path = kuhn_msh(3, 0.15, 2)
    om, X = dof_coords(path)
    s = O.OSim(om, 1, O.CAVITY_MODE, cavity_par(), kin=kin)
    s.set_time(0.5, 0.0, 5.0)
    dt = s.dt
    vi = {O.D3Q4: np.sqrt(3.0) * np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]),
          O.D3Q6: 3.0 * np.vstack([np.eye(3), -np.eye(3)])}[kin]
    ne = om.nelem
    for iv in range(nv):
        a, b = rng.normal(), rng.normal(size=3)
        v = vi[iv]
        f = a + X @ b
        f[10 * ne:] = a + (X[10 * ne:] - 0.5 * dt * v) @ b  # ghosts at t + dt/2
        sol = s.transport(iv, f)
        expect = a + (X[: 10 * ne] - dt * v) @ b
        assert np.abs(sol[: 10 * ne] - expect).max() < 1e-8 * np.abs(expect).max()
        assert np.array_equal(sol[10 * ne:], f[10 * ne:])  # identity ghost rows
        g = rng.normal(size=s.neq)
        x = s.transport(iv, g)
        assert s.transport_residual(iv, g, x) < 1e-13
        nblk, big = s.block_stats(iv)
        assert big == 10 and nblk == ne + 6 * om.nbf
    s.free()
    om.free()


def test_upwind_classes_match_assembled_dependencies():
    """The library's host-side upwind classification (exact reference sign
    test) yields exactly the element couplings of the oracle's assembled
    M1 = I + dt/2 M^-1 F_v."""
    path = kuhn_msh(3, 0.2, 5)
    m = G.Mesh(path=path)
    model = G.Model("cavity_mode", kinetic="d3q4")
    lib = G.load_library()
    cls_p = lib.gdn_upwind_classes(C.byref(m.s), C.byref(model.s))
    ne = m.nelem
    cls = np.ctypeslib.as_array(cls_p, shape=(ne * 16,)).reshape(ne, 4, 4).copy()
    e2e = m.arr("elem2elem", 4 * ne).reshape(ne, 4)
    om = O.OMesh(path)
    s = O.OSim(om, 1, O.CAVITY_MODE, cavity_par(), kin=O.D3Q4)
    s.set_time(0.5, 0.0, 5.0)
    for v in range(4):
        deps = s.m1_dependencies(v)
        for e in range(ne):
            want = set()
            for f in range(4):
                if cls[e, f, v] == 2:
                    want.add(e2e[e, f] if e2e[e, f] >= 0 else -2)
            got = set(int(d) for d in deps[e] if d != -1)
            assert want == got, (v, e, want, got)
    s.free()
    om.free()
    m.free()


def test_relaxation_preserves_moments():
    path = kuhn_msh(2, 0.1, 3)
    om = O.OMesh(path)
    s = O.OSim(om, 1, O.CAVITY_MODE, cavity_par(), kin=O.D3Q4)
    f = rng.normal(size=s.m * s.neq)
    s.set_wn(f)
    s.relax(True)
    g = s.wn()
    w = lambda h: h.reshape(6, 4, s.neq).sum(axis=1)
    assert np.abs(w(g) - w(f)).max() < 1e-13
    # over-relaxation reflects about equilibrium: f_new - feq = (1-om)(f - feq)
    s.set_wn(g)
    s.relax(True)
    h2 = s.wn()
    om_ = 1.999999999999
    feq = (g - (1 - om_) * f) / om_
    assert np.allclose(h2, om_ * feq + (1 - om_) * g, atol=1e-12)
    s.free()
    om.free()


def test_error_norm_of_offset_interpolant():
    """The exact nodal interpolant has zero nodal error; an offset c on one
    component gives sqrt(c^2 * volume) / 6 since the nodal weights of each
    element sum to its volume."""
    path = kuhn_msh(3, 0.2, 8)
    om = O.OMesh(path)
    s = O.OSim(om, 0, O.PLANE_WAVE, plane_wave_par())
    s.init_solution()
    assert s.error_l2() < 1e-14
    w = s.wn().reshape(-1, 6)
    w[: 10 * om.nelem, 2] += 0.25
    s.set_wn(w.ravel())
    assert abs(s.error_l2() - 0.25 / 6) < 1e-14
    s.free()
    om.free()


def test_time_parameters_of_the_benchmark_configurations():
    """dt = hmin cfl / lambda, itermax = floor(tmax/dt), tmax truncated
    (src/simulation/simulation.c:144-164)."""
    m = G.Mesh(kuhn=(8, 0.0, 0))
    sim = G.Simulation(m, G.Model("plane_wave"))
    sim.set_time(1.0, 0.0, 0.5)
    assert sim.s.itermax == 231
    assert abs(sim.s.dt - 4.3147e-3) < 1e-7 and abs(sim.s.tmax - 0.99670) < 1e-5
    sim.free()
    m = G.Mesh(kuhn=(32, 0.0, 0))
    sim = G.Simulation(m, G.Model("cavity_mode", kinetic="d3q4"))
    sim.set_time(2.0, 0.0, 5.0)
    assert sim.s.itermax == 321 and sim.s.neq == 2039808 and sim.s.m == 24
    assert abs(sim.s.dt - 6.228e-3) < 1e-6
    sim.free()


@pytest.mark.parametrize("scheme,jitter", [(0, 0.0), (0, 0.25), (1, 0.1)])
def test_energy_exact_for_quadratic_fields(scheme, jitter):
    """The oracle's energy 1/2 w^T M w against closed forms: for fields that
    are at most quadratic the P2 interpolant is the field itself and M is the
    exact mass, so the energy is 1/2 of the exact integral over the unit
    cube: E_x = 1 gives 1/2, E_y = x gives 1/6, H_z = x y gives 1/18,
    E_z = x^2 gives 1/10 (O(1) off for a lumped mass, a missing 1/2 or a
    wrong det). Kinetic states carry w / nv in every velocity."""
    path = kuhn_msh(3, jitter, 5)
    om, X = dof_coords(path)
    kin = O.D3Q4 if scheme else 0
    s = O.OSim(om, scheme, O.PLANE_WAVE, plane_wave_par(), kin=kin)
    x, y = X[:, 0], X[:, 1]
    fields = [(0, np.ones_like(x), 0.5), (1, x, 1 / 6), (5, x * y, 1 / 18), (2, x * x, 0.1)]
    for q, val, exact in fields:
        w = np.zeros((len(x), 6))
        w[:, q] = val
        if scheme:
            nv = s.m // 6
            f = np.repeat(w.T[:, None, :], nv, axis=1) / nv  # [q][v][ipg]
            s.set_wn(f.reshape(-1))
        else:
            s.set_wn(w.reshape(-1))
        e = s.energy()
        assert abs(e - exact) <= 1e-13, (q, e, exact)
    s.free()
    om.free()


# ------------------------------------------------------------- permittivity
from helpers import block_eps, gaussian_par, plane_wave_eps_par  # noqa: E402


@pytest.mark.parametrize("scheme", [0, 1])
def test_uniform_eps_plane_wave_converges(scheme):
    """eps = 4 everywhere: the exact solution is the plane wave of speed
    1/2 with H = 2 (k/|k|) x E. The oracle's RK3 (eps on the E rows of
    M^-1) and kinetic scheme (conserved (D, B), equilibrium flux on
    (D/eps, B)) converge to it: the error falls by >= 3.5x from 3^3 to 6^3
    cubes and is below 4e-3 (RK3, CFL 0.5) or 3e-2 (kinetic, cfl 5; its
    eps = 1 error on this case is 1.7e-2) at 6^3 (O(1) for eps on the wrong
    rows, for the transport slowed by eps or an equilibrium flux taken on
    D: the wave then travels at the wrong speed)."""
    k = 2 * np.pi * np.array([1.0, 0.0, 0.0])
    par = plane_wave_eps_par(k, (0, 1, 0), 4.0)
    errs = []
    for n in (3, 6):
        om = O.OMesh(kuhn_msh(n, 0.0, 0))
        s = O.OSim(om, scheme, O.PLANE_WAVE, par, kin=O.D3Q4 if scheme else 0,
                   eps=np.full(om.nelem, 4.0))
        s.set_time(0.3, 0.0, 5.0 if scheme else 0.5)
        s.solve()
        errs.append(s.error_l2())
        s.free()
        om.free()
    print(f"uniform eps=4, scheme {scheme}: L2 error 3^3 {errs[0]:.3e}, 6^3 {errs[1]:.3e}")
    assert errs[1] < (3e-2 if scheme else 4e-3) and errs[0] / errs[1] >= 3.5


def test_rk3_eps_energy_identity():
    """The RK3 operator with a piecewise permittivity keeps the energy
    identity <w, dw>_eps = <w, dw_1>_1 = w^T F w (M_eps dw_eps = F w =
    M_1 dw_1) with <u, v>_eps = sum_e det_e (sum_{q<3} eps_e u_q^T M v_q +
    sum_{q>=3} u_q^T M v_q): it holds only if the 1/eps sits on the E rows
    of every element; and w^T F w >= 0 (upwind dissipation, zero ghosts)."""
    path = kuhn_msh(3, 0.2, 8)
    om, X = dof_coords(path)
    eps = block_eps(path, 2, 5)
    ne = om.nelem
    T = O.tables()
    det = np.array([np.linalg.det(np.stack([X[10 * e + v] - X[10 * e] for v in (1, 2, 3)]))
                    for e in range(ne)])
    w = rng.normal(size=(len(X), 6))
    w[10 * ne:] = 0
    out = {}
    for tag, ep in (("eps", eps), ("one", np.ones(ne))):
        s = O.OSim(om, 0, O.PLANE_WAVE, plane_wave_par(), eps=ep)
        out[tag] = s.rk3_operator(w.ravel()).reshape(-1, 6)
        s.free()

    def inner(u, v, ep):
        tot = 0.0
        for e in range(ne):
            ue, ve = u[10 * e:10 * e + 10], v[10 * e:10 * e + 10]
            for q in range(6):
                tot += det[e] * (ep[e] if q < 3 else 1.0) * ue[:, q] @ T["mass"] @ ve[:, q]
        return tot

    a, b = inner(w, out["eps"], eps), inner(w, out["one"], np.ones(ne))
    print(f"<w, dw>_eps = {a:.15e}, <w, dw_1>_1 = {b:.15e}")
    assert abs(a - b) <= 1e-12 * abs(b)
    assert b > 0  # w^T F w > 0: the update subtracts dt dw, the upwind flux dissipates
    assert not np.allclose(out["eps"], out["one"])
    om.free()


def test_kinetic_eps_equilibrium_moments():
    """The relaxation with a permittivity: from any state f, the
    equilibrium it relaxes towards (feq = (f' - (1 - omega) f) / omega) has
    sum_v feq = w = sum_v f (the conserved (D, B)) and
    sum_v v feq = A(e_k) (D / eps, B), at every element DoF."""
    path = kuhn_msh(2, 0.1, 4)
    om = O.OMesh(path)
    ne = om.nelem
    eps = np.linspace(1.0, 4.0, ne)
    s = O.OSim(om, 1, O.PLANE_WAVE, plane_wave_par(), kin=O.D3Q4, eps=eps)
    f = rng.normal(size=s.m * s.neq)
    s.set_wn(f)
    s.relax(with_ghosts=False)
    f1 = s.wn()
    om_ = 1.999999999999
    feq = ((f1 - (1 - om_) * f) / om_).reshape(6, 4, s.neq)
    F = f.reshape(6, 4, s.neq)
    vi = np.sqrt(3.0) * np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]])
    for ipg in range(0, 10 * ne, 7):
        w = F[:, :, ipg].sum(axis=1)
        wt = w.copy()
        wt[:3] /= eps[ipg // 10]
        assert np.allclose(feq[:, :, ipg].sum(axis=1), w, atol=1e-9)
        q = np.stack([A(np.eye(3)[k]) @ wt for k in range(3)], axis=1)
        assert np.allclose(feq[:, :, ipg] @ vi, q, atol=1e-9)
    s.free()
    om.free()


def test_kinetic_and_rk3_agree_with_piecewise_eps():
    """Both schemes discretise eps dE/dt = curl H, dH/dt = -curl E with a
    piecewise permittivity (eps in [1, 4] per 2^3 block), RK3 on (E, H) and
    the kinetic scheme on the conserved (D, B). At the material interfaces
    the kinetic upwind transport of f_eq(D, B) adds a dissipation of the jump
    of D (tangential D jumps where E does not), so the two differ there by
    more than the discretisation error of either (measured: 0.37 relative at
    8^3); the kinetic scheme with the permittivity still lies closer to RK3
    than without it (0.58), which a permittivity dropped from or misplaced in
    its equilibrium would not."""
    path = kuhn_msh(8, 0.0, 0)
    eps = block_eps(path, 2, 7)
    par = gaussian_par(sigma=0.2)
    om = O.OMesh(path)
    ne = om.nelem
    res = {}
    for tag, scheme, ep in (("rk3", 0, eps), ("kin", 1, eps), ("kin1", 1, np.ones(ne))):
        s = O.OSim(om, scheme, 6, par, kin=O.D3Q4 if scheme else 0, eps=ep)
        s.set_time(0.3, 0.0, 2.0 if scheme else 0.5)
        s.solve()
        w = s.wn()
        if scheme:
            E = w.reshape(6, 4, s.neq).sum(axis=1)[:3, : 10 * ne] / np.repeat(ep, 10)[None, :]
        else:
            E = w.reshape(-1, 6)[: 10 * ne, :3].T
        res[tag] = E
        s.free()
    om.free()
    d = np.linalg.norm(res["kin"] - res["rk3"]) / np.linalg.norm(res["rk3"])
    d1 = np.linalg.norm(res["kin1"] - res["rk3"]) / np.linalg.norm(res["rk3"])
    print(f"piecewise eps, E at t=0.3: kinetic vs RK3 {d:.3e}; kinetic with eps=1 vs RK3 {d1:.3e}")
    assert d < 0.8 * d1


def test_unfactored_assembly_gives_the_same_residual():
    """factor=False assembles M1 and M2 without factoring M1: the transport
    residual is the factored simulation's, and a perturbed solution shows."""
    path = kuhn_msh(3, 0.1, 4)
    om = O.OMesh(path)
    a = O.OSim(om, 1, O.CAVITY_MODE, cavity_par(), kin=O.D3Q4)
    b = O.OSim(om, 1, O.CAVITY_MODE, cavity_par(), kin=O.D3Q4, factor=False)
    a.set_time(0.3, 0.0, 5.0)
    b.set_time(0.3, 0.0, 5.0)
    f = np.random.default_rng(8).normal(size=a.neq)
    x = a.transport(1, f)
    ra, rb = a.transport_residual(1, f, x), b.transport_residual(1, f, x)
    assert ra == rb and rb < 1e-13
    x[17] += 1e-3
    assert b.transport_residual(1, f, x) > 1e-7
    a.free()
    b.free()
    om.free()
