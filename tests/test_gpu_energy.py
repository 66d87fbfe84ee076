"""GPU energy reduction 1/2 w^T M w (gdh_energy on a context,
gdh_energy_dev on a state held in device memory by the caller) against the
oracle's assembled-mass energy (oracle/oracle.c orc_energy, itself pinned by
closed forms in tests/test_cpu_oracle.py): relative difference <= 1e-12
after whole RK3 and relaxation solves; closed forms on the device."""
import numpy as np
import pytest

import gdon3d as G
import oracle as O
from helpers import HipRuntime, cavity_par, kuhn_msh, plane_wave_par

TOL = 1e-12


def det_of(path):
    mesh = G.Mesh(path=path)
    det = np.zeros(mesh.nelem)
    d = np.zeros(1)
    lib = G.load_library()
    for e in range(mesh.nelem):
        lib.mesh_get_elem_det(G.C.byref(mesh.s), e, G.dptr(d))
        det[e] = d[0]
    mesh.free()
    return det


@pytest.mark.gpu
@pytest.mark.parametrize("kin,precision", [(None, "fp64"), ("d3q4", "fp64"), ("d3q6", "fp64"),
                                           ("d3q4", "fp32")])
def test_energy_after_solves_vs_oracle(kin, precision):
    path = kuhn_msh(4, 0.15, 3)
    if kin:
        G.set_cavity_mode(1, 1, 1, (1, -1, 0))
        imposed, okind, par, cfl, tend = "cavity_mode", O.CAVITY_MODE, cavity_par(), 5.0, 0.3
    else:
        G.set_plane_wave((2 * np.pi, 0, 0), (0, 1, 0))
        imposed, okind, par, cfl, tend = "plane_wave", O.PLANE_WAVE, plane_wave_par(), 0.5, 0.2
    sim = G.Simulation(G.Mesh(path=path), G.Model(imposed, kinetic=kin))
    sim.set_time(tend, 0.0, cfl)
    st = G.Stepper(sim, precision=precision)
    e0 = st.energy()
    st.advance(sim.s.itermax)
    st.sync()
    e1 = st.energy()
    st.free()
    sim.free()
    om = O.OMesh(path)
    s = O.OSim(om, 1 if kin else 0, okind, par,
               kin={None: 0, "d3q4": O.D3Q4, "d3q6": O.D3Q6}[kin])
    s.set_time(tend, 0.0, cfl)
    s.init_solution()
    o0 = s.energy()
    s.solve()
    o1 = s.energy()
    s.free()
    om.free()
    tol = TOL if precision == "fp64" else 2e-6
    print(f"energy {kin} {precision}: GPU {e0:.15e} -> {e1:.15e}, oracle {o0:.15e} -> "
          f"{o1:.15e}", flush=True)
    assert abs(e0 - o0) <= tol * o0 and abs(e1 - o1) <= tol * o1


@pytest.mark.gpu
@pytest.mark.parametrize("kin", [None, "d3q4"])
def test_energy_closed_form_on_device(kin):
    """Uploaded quadratic fields: E_y = x gives 1/6 and E_z = x^2 + H_x = 1
    gives 1/10 + 1/2 over the unit cube (exact P2 interpolation and mass)."""
    path = kuhn_msh(3, 0.2, 9)
    G.set_plane_wave((2 * np.pi, 0, 0), (0, 1, 0))
    sim = G.Simulation(G.Mesh(path=path), G.Model("plane_wave", kinetic=kin))
    sim.set_time(0.1, 0.0, 0.5)
    st = G.Stepper(sim)
    om = O.OMesh(path)
    a = om.arrays()
    X = a["node"].reshape(-1, 3)[a["elem2qnode"]]
    om.free()
    x = X[:, 0]
    for fill, exact in (({1: x}, 1 / 6), ({2: x * x, 3: np.ones_like(x)}, 0.1 + 0.5)):
        w = np.zeros((len(x), 6))
        for q, v in fill.items():
            w[:, q] = v
        if kin:
            nv = sim.s.m // 6
            wn = (np.repeat(w.T[:, None, :], nv, axis=1) / nv).reshape(-1)
        else:
            wn = w.reshape(-1)
        st.upload(wn)
        e = st.energy()
        print(f"closed form {kin}: {e:.16f} vs {exact:.16f}", flush=True)
        assert abs(e - exact) <= 1e-13
    st.free()
    sim.free()


@pytest.mark.gpu
@pytest.mark.parametrize("kin", [None, "d3q4"])
def test_energy_dev_pointer_entry(kin):
    """gdh_energy_dev on a state the caller holds in device memory (the
    reference layout of simu->wn) gives the context's gdh_energy."""
    path = kuhn_msh(5, 0.1, 2)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    sim = G.Simulation(G.Mesh(path=path), G.Model("cavity_mode", kinetic=kin))
    sim.set_time(0.2, 0.0, 5.0 if kin else 0.5)
    st = G.Stepper(sim)
    st.advance(3)
    st.sync()
    e_ctx = st.energy()
    wn = st.download()
    ne, nb = sim.mesh.s.nbelems, sim.mesh.s.nboundaryfaces
    nv = sim.s.m // 6 if kin else 0
    st.free()
    det = det_of(path)
    hip = HipRuntime()
    dw, dd, do = hip.to_device(wn), hip.to_device(det), hip.alloc(8)
    rc = G.load_library().gdh_energy_dev(dw, nv, ne, nb, dd, do, None)
    e_dev = hip.to_host(do, 1)[0]
    for p in (dw, dd, do):
        hip.free(p)
    sim.free()
    print(f"energy {kin}: context {e_ctx:.16e}, device pointers {e_dev:.16e}", flush=True)
    assert rc == 0
    assert abs(e_dev - e_ctx) <= TOL * e_ctx
