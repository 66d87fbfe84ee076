"""FP32 precision form (dg_stepper_create_precision(simu, GDN_PRECISION_FP32),
gdh_create_ex(..., GDH_FP32)): fields stored in binary32 and the RK3 flux /
volume arithmetic in binary32; geometry, transport block factors, transport
and relaxation arithmetic and the L2 reductions in binary64.

Reference for each case: the plain FP64 GPU form on the same mesh, which the
other GPU tests pin to the oracle at relative L2 <= 1e-12, and the exact
solution.

Tolerances (binary32 unit roundoff 6e-8; 100-300 steps):
* FP32 state vs FP64 state: relative L2 <= 2e-5 after the whole solve;
* FP32 L2 error vs the exact solution within 1 % (+1e-7 absolute) of the
  FP64 error: the rounding stays below the discretisation error.
The FP32 persistent sweep equals the FP32 level-by-level sweep bit for bit.
"""
import numpy as np
import pytest

import gdon3d as G
from helpers import kuhn_msh, rel_l2

REL_STATE = 2e-5


def solve(path, imposed, kinetic, tend, cfl, precision):
    mesh = G.Mesh(path=path)
    sim = G.Simulation(mesh, G.Model(imposed, kinetic=kinetic))
    sim.set_time(tend, 0.0, cfl)
    st = G.Stepper(sim, precision=precision)
    st.advance(sim.s.itermax)
    st.sync()
    err = st.error_l2()
    st.finish()
    out = dict(wn=sim.wn(), err=err, t=st.t, steps=sim.s.itermax)
    st.free()
    sim.free()
    return out


def oblique_wave():
    k = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)
    G.set_plane_wave(k, (0.0, 0.0, 1.0))


def check(a, b, label):
    d = rel_l2(a["wn"], b["wn"])
    print(f"{label}: steps {a['steps']} rel state diff {d:.3e} "
          f"err fp32 {a['err']:.6e} fp64 {b['err']:.6e}", flush=True)
    assert a["steps"] == b["steps"] and abs(a["t"] - b["t"]) < 1e-12
    assert d <= REL_STATE
    assert abs(a["err"] - b["err"]) <= 0.01 * b["err"] + 1e-7


@pytest.mark.gpu
def test_fp32_rk3_oblique_plane_wave_jittered():
    """RK3, k = 2 pi (1,1,0)/sqrt 2, E along z, +-10 % jitter (seed 17),
    CFL 0.5, t = 1: the precision-form workload at 8^3."""
    print("fp32 rk3 8^3 jitter 0.1", flush=True)
    path = kuhn_msh(8, 0.1, 17)
    oblique_wave()
    a = solve(path, "plane_wave", None, 1.0, 0.5, "fp32")
    b = solve(path, "plane_wave", None, 1.0, 0.5, "fp64")
    check(a, b, "rk3")


@pytest.mark.gpu
def test_fp32_relaxation_oblique_plane_wave_jittered():
    """Relaxation D3Q4, the same wave, +-10 % jitter, cfl 5, t = 0.5."""
    print("fp32 relaxation 6^3 jitter 0.1", flush=True)
    path = kuhn_msh(6, 0.1, 17)
    oblique_wave()
    a = solve(path, "plane_wave", "d3q4", 0.5, 5.0, "fp32")
    b = solve(path, "plane_wave", "d3q4", 0.5, 5.0, "fp64")
    check(a, b, "relax plane wave")


@pytest.mark.gpu
def test_fp32_relaxation_cavity_d3q6():
    print("fp32 relaxation d3q6 cavity 4^3", flush=True)
    path = kuhn_msh(4, 0.0, 3)
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    a = solve(path, "cavity_mode", "d3q6", 0.4, 3.0, "fp32")
    b = solve(path, "cavity_mode", "d3q6", 0.4, 3.0, "fp64")
    check(a, b, "relax cavity d3q6")


@pytest.mark.gpu
@pytest.mark.parametrize("poll", ["sentinel", "full"])
def test_fp32_persistent_sweep_equals_level_sweep(poll, monkeypatch):
    path = kuhn_msh(6, 0.2, 31)
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    monkeypatch.setenv("GDN_SWEEP_POLL", poll)
    out = {}
    for mode in ("level", "persistent"):
        monkeypatch.setenv("GDN_SWEEP", mode)
        out[mode] = solve(path, "cavity_mode", "d3q4", 0.3, 5.0, "fp32")["wn"]
    assert np.array_equal(out["level"], out["persistent"])
