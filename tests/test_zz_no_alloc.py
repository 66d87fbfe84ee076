"""gdh_energy_dev_noalloc, the energy entry point that allocates nothing
(its partial sums in a fixed device array), against gdh_energy_dev (the same
kernels with a buffer allocated per call): bit for bit, on two calls in a
row, the second on a smaller state, after any other test; and against the
context's gdh_energy to 1e-12."""
import numpy as np
import pytest

import gdon3d as G
from helpers import HipRuntime, cavity_par, kuhn_msh  # noqa: F401


def state_of(n, kin):
    """A kinetic (or RK3) state after a few steps, its determinants and the
    context's energy."""
    path = kuhn_msh(n, 0.1, 4)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    sim = G.Simulation(G.Mesh(path=path), G.Model("cavity_mode", kinetic=kin))
    sim.set_time(0.2, 0.0, 5.0 if kin else 0.5)
    st = G.Stepper(sim)
    st.advance(2)
    st.sync()
    e_ctx = st.energy()
    wn = st.download()
    ne, nb = sim.mesh.s.nbelems, sim.mesh.s.nboundaryfaces
    nv = sim.s.m // 6 if kin else 0
    lib = G.load_library()
    det = np.zeros(ne)
    d = np.zeros(1)
    for e in range(ne):
        lib.mesh_get_elem_det(G.C.byref(sim.mesh.s), e, G.dptr(d))
        det[e] = d[0]
    st.free()
    sim.free()
    return wn, det, ne, nb, nv, e_ctx


@pytest.mark.gpu
@pytest.mark.parametrize("kin", ["d3q4", None])
def test_energy_noalloc_equals_allocating_entry(kin):
    lib = G.load_library()
    hip = HipRuntime()
    for n in (12, 5):  # the second call on a smaller state
        wn, det, ne, nb, nv, e_ctx = state_of(n, kin)
        print(f"no-alloc energy: {kin} n={n} ne={ne} nv={nv}", flush=True)
        dw, dd = hip.to_device(wn), hip.to_device(det)
        out_a, out_b = hip.alloc(8), hip.alloc(8)
        assert lib.gdh_energy_dev(dw, nv, ne, nb, dd, out_a, None) == 0
        assert lib.gdh_energy_dev_noalloc(dw, nv, ne, nb, dd, out_b, None) == 0
        ea, eb = hip.to_host(out_a, 1)[0], hip.to_host(out_b, 1)[0]
        for p in (dw, dd, out_a, out_b):
            hip.free(p)
        print(f"  allocating {ea:.17e}, no-alloc {eb:.17e}, context {e_ctx:.17e}", flush=True)
        assert ea == eb
        assert abs(eb - e_ctx) <= 1e-12 * e_ctx


@pytest.mark.gpu
def test_energy_noalloc_rejects_invalid_arguments():
    lib = G.load_library()
    hip = HipRuntime()
    out = hip.alloc(8)
    assert lib.gdh_energy_dev_noalloc(None, 4, 10, 0, out, out, None) == -1
    assert lib.gdh_energy_dev_noalloc(out, 4, -1, 0, out, out, None) == -1
    assert lib.gdh_energy_dev_noalloc(out, 99, 1, 0, out, out, None) == -1
    hip.free(out)
