"""The reference's standalone dg_* entry points on the device
(src/dg/dg.c:32-147): dg_init_sol_macro, dg_update_boundary and
dg_apply_relaxation, called on a gdn_simulation one after another as a
reference-style driver does, against the oracle's init_sol, update_boundary
and relax (boundary DoFs included) on the same states: relative L2 <= 1e-13.

The calls share one device context kept for the simulation; the state is
uploaded at every call, so a state changed on the host between calls is the
one used, and a simulation on another mesh gets a context of its own.
"""
import ctypes as C

import numpy as np
import pytest

import gdon3d as G
import oracle as O
from helpers import kuhn_msh, plane_wave_par, rel_l2

TOL = 1e-13
K = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2)

CASES = {
    "rk3": (None, "feq", 0, False),
    "d3q4": ("d3q4", "feq", O.D3Q4, False),
    "d3q6": ("d3q6", "feq", O.D3Q6, False),
    "d3q4-projector": ("d3q4", "projector", O.D3Q4, True),
    "d3q6-projector": ("d3q6", "projector", O.D3Q6, True),
}


def run_sequence(path, case, states):
    """The entry points on the device and in the oracle: init, then for
    each (time, seed) a seeded random state set on the host, an update at
    that time and, for the kinetic scheme, a relaxation. Returns the pairs
    (device, oracle) of every result."""
    kin, bstate, okin, proj = CASES[case]
    lib = G.load_library()
    G.set_plane_wave(K, (0, 0, 1))
    sim = G.Simulation(G.Mesh(path=path),
                       G.Model("plane_wave", kinetic=kin, boundary_state=bstate))
    om = O.OMesh(path)
    osim = O.OSim(om, 0 if kin is None else 1, O.PLANE_WAVE, plane_wave_par(K, (0, 0, 1)),
                  kin=okin, fr_projector=proj)
    assert sim.s.wlen == osim.m * osim.neq
    out = []
    lib.dg_init_sol_macro(C.byref(sim.s))
    osim.init_solution()
    out.append(("init", sim.wn(), osim.wn()))
    for t, seed in states:
        w = np.random.default_rng(seed).normal(size=sim.s.wlen)
        sim.set_wn(w)
        osim.set_wn(w)
        lib.dg_update_boundary(C.byref(sim.s), t)
        osim.update_boundary(t)
        out.append((f"update t={t}", sim.wn(), osim.wn()))
        if kin is not None:
            lib.dg_apply_relaxation(C.byref(sim.s))
            osim.relax(with_ghosts=True)
            out.append(("relax", sim.wn(), osim.wn()))
    sim.free()
    osim.free()
    om.free()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(CASES))
def test_entry_points_vs_oracle(case):
    path = kuhn_msh(3, 0.15, 29)
    for what, dev, ref in run_sequence(path, case, [(0.2, 31), (0.37, 32)]):
        r = rel_l2(dev, ref)
        print(f"  {case} {what}: rel L2 vs oracle {r:.3e}", flush=True)
        assert r <= TOL, what
    # the kept context follows the simulation: another mesh, the same calls
    path2 = kuhn_msh(4, 0.1, 30)
    for what, dev, ref in run_sequence(path2, case, [(0.11, 33)]):
        r = rel_l2(dev, ref)
        print(f"  {case} second mesh {what}: rel L2 vs oracle {r:.3e}", flush=True)
        assert r <= TOL, what
