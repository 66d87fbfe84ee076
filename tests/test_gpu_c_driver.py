"""A project-written C program (tests/c/solve_driver.c) making the
reference's call sequence through include/ and libgdon3d_hip.so —
mesh_read_from_msh22_file, mesh_build_connectivity, simulation_init,
simulation_set_time_parameters, dg_solve, simulation_error_l2,
simulation_dump_info, simulation_free (src/utils/convergence.c:43-98) — on
the reference's own mesh t4_cube_8.msh, RK3 and the relaxation scheme
(D3Q4, D3Q6), against the oracle: the state after the solve to relative
L2 <= 1e-12, the same step count, the L2 error to 1e-12 relative.

The test process keeps a device context of its own while the driver runs:
a second process on the device, where the runtime refuses cooperative
launches (measured on MI355X: "the operation cannot be performed in the
present state"); the library then runs the item sweep over its queue, and
the driver's results are checked all the same."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import plane_wave_par, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "c", "solve_driver")
MESH = os.path.join(HERE, "golden", "t4_cube_8.msh")
K = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme,tmax,cfl", [("rk3", 0.1, 0.5), ("d3q4", 0.3, 5.0),
                                             ("d3q6", 0.2, 3.0)])
def test_c_driver_call_sequence_vs_oracle(scheme, tmax, cfl, tmp_path):
    assert os.path.exists(DRIVER), "build it first: make -C tests/c"
    import gdon3d as G
    G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
    held = G.Simulation(G.Mesh(kuhn=(2, 0.0, 0)), G.Model("cavity_mode", kinetic="d3q4"))
    held.set_time(0.1, 0.0, 5.0)
    held_st = G.Stepper(held)  # this process's device context, alive during the run
    out = str(tmp_path / "wn.bin")
    print(f"C driver {scheme} tmax {tmax} cfl {cfl}", flush=True)
    r = subprocess.run([DRIVER, MESH, scheme, str(tmax), str(cfl), out], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout[-600:], r.stderr[-600:], flush=True)
    held_st.free()
    held.free()
    assert r.returncode == 0
    print(f"  cooperative launch refused in the driver: {'refused a cooperative' in r.stderr}",
          flush=True)
    itermax, err = r.stdout.strip().splitlines()[-1].split()
    wn = np.fromfile(out, dtype=np.float64)
    om = O.OMesh(MESH)
    kin = {"rk3": 0, "d3q4": O.D3Q4, "d3q6": O.D3Q6}[scheme]
    s = O.OSim(om, 0 if scheme == "rk3" else 1, O.PLANE_WAVE, plane_wave_par(K, (0, 0, 1)),
               kin=kin)
    s.set_time(tmax, 0.0, cfl)
    s.solve()
    ref, oerr, oit = s.wn(), s.error_l2(), s.itermax
    s.free()
    om.free()
    rd = rel_l2(wn, ref)
    print(f"  steps {itermax} (oracle {oit}), error {float(err):.6e} (oracle {oerr:.6e}), "
          f"state rel L2 vs oracle {rd:.3e}", flush=True)
    assert int(itermax) == oit
    assert rd <= 1e-12
    assert abs(float(err) - oerr) <= 1e-12 * oerr
