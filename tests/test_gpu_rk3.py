"""GPU parity of the RK3 solve (dg_solve_rk3 through the C ABI) against the
oracle's assembled-matrix restatement of src/dg/explicit.c:39-117.

Tolerance: relative L2 <= 1e-12 on the full state wn (FP64)."""
import numpy as np
import pytest

import gdon3d as G
import oracle as O
from helpers import kuhn_msh, plane_wave_par, cavity_par, rel_l2

TOL = 1e-12


def run_gpu(path, imposed, tend, cfl, boundary="upwind", callback=None):
    mesh = G.Mesh(path=path)
    model = G.Model(imposed, boundary=boundary, callback=callback)
    sim = G.Simulation(mesh, model)
    sim.set_time(tend, 0.0, cfl)
    sim.solve()
    out = dict(wn=sim.wn(), err=sim.error_l2(), itermax=sim.s.itermax, t=sim.s.t,
               tmax=sim.s.tmax)
    sim.free()
    return out


def run_oracle(path, kind, par, tend, cfl, bflux=O.UPWIND):
    om = O.OMesh(path)
    s = O.OSim(om, 0, kind, par, bflux=bflux)
    s.set_time(tend, 0.0, cfl)
    s.solve()
    out = dict(wn=s.wn(), err=s.error_l2(), itermax=s.itermax, t=s.t)
    s.free()
    om.free()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n,jitter,tend", [(3, 0.0, 0.25), (4, 0.15, 0.3), (8, 0.0, 1.0)])
def test_rk3_plane_wave(n, jitter, tend):
    """Ey = Hz = sin(2 pi (x - t)), CFL 0.5; n = 8, t_end = 1 is the
    3072-tet plane-wave configuration (231 steps)."""
    print(f"rk3 plane wave n={n} jitter={jitter} tend={tend}", flush=True)
    path = kuhn_msh(n, jitter, 17)
    G.set_plane_wave((2 * np.pi, 0, 0), (0, 1, 0))
    g = run_gpu(path, "plane_wave", tend, 0.5)
    o = run_oracle(path, O.PLANE_WAVE, plane_wave_par(), tend, 0.5)
    assert g["itermax"] == o["itermax"]
    assert abs(g["t"] - o["t"]) < 1e-12
    assert rel_l2(g["wn"], o["wn"]) <= TOL
    assert abs(g["err"] - o["err"]) <= TOL * o["err"]
    if n == 8:
        assert g["itermax"] == 231
        assert g["err"] < 5e-3  # P2 on h = 1/8 against the exact wave


@pytest.mark.gpu
def test_rk3_oblique_wave_jittered():
    path = kuhn_msh(4, 0.1, 5)
    k = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2)
    G.set_plane_wave(k, (0, 0, 1), 0.3)
    g = run_gpu(path, "plane_wave", 0.2, 0.5)
    o = run_oracle(path, O.PLANE_WAVE, plane_wave_par(k, (0, 0, 1), 0.3), 0.2, 0.5)
    assert rel_l2(g["wn"], o["wn"]) <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("imposed,kind", [("macro_cos", O.MACRO_COS),
                                          ("macro_cavity", O.CAVITY_TM),
                                          ("stationary", O.STATIONARY)])
def test_rk3_reference_solutions(imposed, kind):
    path = kuhn_msh(3, 0.1, 3)
    g = run_gpu(path, imposed, 0.15, 0.4)
    o = run_oracle(path, kind, None, 0.15, 0.4)
    assert rel_l2(g["wn"], o["wn"]) <= TOL


@pytest.mark.gpu
def test_rk3_metal_boundary_cavity():
    """PEC boundary flux (maxwell_num_flux_boundary_metal) on the (1,1,1)
    cavity mode."""
    path = kuhn_msh(4, 0.0, 0)
    G.set_cavity_mode(1, 1, 1, (1, -1, 0))
    g = run_gpu(path, "cavity_mode", 0.3, 0.5, boundary="metal")
    o = run_oracle(path, O.CAVITY_MODE, cavity_par(), 0.3, 0.5, bflux=O.METAL)
    assert rel_l2(g["wn"], o["wn"]) <= TOL


@pytest.mark.gpu
def test_rk3_host_callback_ghosts():
    """An imposed-data callback unknown to the library: ghosts are evaluated
    on the host each stage and uploaded."""
    def cb(x, t, w):
        s = np.sin(2 * np.pi * (x[0] - t))
        for q in range(6):
            w[q] = 0.0
        w[1] = s
        w[5] = s

    path = kuhn_msh(3, 0.0, 0)
    g = run_gpu(path, None, 0.2, 0.5, callback=cb)
    o = run_oracle(path, O.PLANE_WAVE, plane_wave_par(), 0.2, 0.5)
    assert rel_l2(g["wn"], o["wn"]) <= TOL


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("boundary", ["upwind", "metal"])
def test_rk3_tiled_stage_equals_global_stage(precision, boundary, monkeypatch):
    """The LDS-tiled RK3 stages with one lane per element (GDN_RK3=tile, the
    binary32 default) and two (GDN_RK3=pair, the binary64 default) and the
    stage over the global state (GDN_RK3=soa) compute the same sums; the compiler fuses and orders the
    floating-point operations of the two kernels differently, so the states
    after a whole solve agree to rounding: relative L2 <= 1e-13 (FP64),
    <= 1e-5 (FP32). 12^3 x 6 = 10368 tets spans 162 tiles of 64 elements,
    with neighbours inside and outside the tile and boundary faces."""
    path = kuhn_msh(12, 0.1, 9)
    if boundary == "metal":
        G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
        imposed = "cavity_mode"
    else:
        G.set_plane_wave((2 * np.pi, 0, 0), (0, 1, 0))
        imposed = "plane_wave"
    out = {}
    for k in ("soa", "tile", "pair"):
        monkeypatch.setenv("GDN_RK3", k)
        mesh = G.Mesh(path=path)
        sim = G.Simulation(mesh, G.Model(imposed, boundary=boundary))
        sim.set_time(0.05, 0.0, 0.5)
        st = G.Stepper(sim, precision=precision)
        st.advance(sim.s.itermax)
        st.sync()
        st.finish()
        out[k] = sim.wn()
        st.free()
        sim.free()
    for k in ("tile", "pair"):
        d = rel_l2(out[k], out["soa"])
        print(f"{k} vs global RK3 stage, {precision} {boundary}: rel L2 {d:.3e}", flush=True)
        assert d <= (1e-13 if precision == "fp64" else 1e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rk3_block_sizes_bit_identical(precision, monkeypatch):
    """Each RK3 stage kernel at every block size it has (GDN_RK3_BLOCK:
    elements per block; the LDS-tile kernels stage that many columns, the
    global-state kernel takes that many threads) performs the same arithmetic
    per element: only whether a face neighbour's values come from LDS or from
    global memory changes, not the values. So every block size gives the same
    state bit for bit after a whole solve (12^3 x 6 tets, jittered, upwind
    boundary: tiles of 32 to 256 columns, neighbours inside and outside), and
    the one against the oracle holds for all. The context reports the block
    size it runs."""
    path = kuhn_msh(12, 0.1, 9)
    G.set_plane_wave((2 * np.pi, 0, 0), (0, 1, 0))
    sizes = {"pair": (32, 64, 128), "tile": (64, 128) + ((256,) if precision == "fp32" else ()),
             "soa": (64, 128, 256)}
    for kern, blocks in sizes.items():
        monkeypatch.setenv("GDN_RK3", kern)
        out = {}
        for bs in blocks:
            monkeypatch.setenv("GDN_RK3_BLOCK", str(bs))
            mesh = G.Mesh(path=path)
            sim = G.Simulation(mesh, G.Model("plane_wave"))
            sim.set_time(0.05, 0.0, 0.5)
            st = G.Stepper(sim, precision=precision)
            info = G.rk3_info(st)
            assert info["kernel"] == kern and info["block"] == bs, info
            st.advance(sim.s.itermax)
            st.sync()
            st.finish()
            out[bs] = sim.wn()
            st.free()
            sim.free()
        ref = out[blocks[0]]
        for bs in blocks[1:]:
            print(f"{kern} {precision}: block {bs} vs {blocks[0]}: max |diff| = "
                  f"{np.abs(out[bs] - ref).max():.3e}", flush=True)
            assert np.array_equal(out[bs], ref)
    o = run_oracle(path, O.PLANE_WAVE, plane_wave_par(), 0.05, 0.5)
    d = rel_l2(out[blocks[0]], o["wn"])
    print(f"soa {precision} vs oracle: rel L2 {d:.3e}", flush=True)
    assert d <= (TOL if precision == "fp64" else 2e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_rk3_k2_stream_hints_bit_identical(precision, monkeypatch):
    """The RK3 stage kernels load and store k2 with non-temporal hints
    (GDN_RK3_K2=nt, the default for one mesh) or ordinary accesses
    (GDN_RK3_K2=plain, which the batch autotuner may pick): a cache hint
    changes no value, so every stage kernel gives the same state bit for bit
    after a whole solve (6^3 x 6 tets, jittered, upwind boundary). The
    context reports the kernel and hint it runs, so a library that ignored
    the variables would fail here."""
    path = kuhn_msh(6, 0.1, 3)
    G.set_plane_wave((2 * np.pi, 0, 0), (0, 1, 0))
    for kern in ("soa", "tile", "pair"):
        monkeypatch.setenv("GDN_RK3", kern)
        out = {}
        for hint in ("nt", "plain"):
            monkeypatch.setenv("GDN_RK3_K2", hint)
            mesh = G.Mesh(path=path)
            sim = G.Simulation(mesh, G.Model("plane_wave"))
            sim.set_time(0.1, 0.0, 0.5)
            st = G.Stepper(sim, precision=precision)
            info = G.rk3_info(st)
            assert info["kernel"] == kern
            assert info["k2_nontemporal"] == (hint == "nt")
            st.advance(sim.s.itermax)
            st.sync()
            st.finish()
            out[hint] = sim.wn()
            st.free()
            sim.free()
        print(f"{kern} {precision}: max |nt - plain| = "
              f"{np.abs(out['nt'] - out['plain']).max():.3e}", flush=True)
        assert np.array_equal(out["nt"], out["plain"])
