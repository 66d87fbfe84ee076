#!/usr/bin/env python3
"""Benchmark of the CFL-less relaxation hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): Maxwell,
D3Q4 kinetic relaxation scheme, FP64, resonant mode (1,1,1) of the perfectly
conducting unit cube as exact solution and ghost data, Kuhn mesh 32^3 x 6 =
196,608 P2 tetrahedra, cfl = 5 (10x the explicit 0.5, lambda = sqrt 3),
t_end = 2.0 -> 321 steps. A step is one full time step: ghosts at t + dt/2,
theta = 1/2 transport of the 24 micro-fields (upwind block sweep) and the
node-local relaxation.

value = DoF-updates/s = (24 * 10 * nelem) * K / (wall time of K steps), with
the state resident in HBM (no host transfer inside the timed region).
roofline: the transport sweep (dominant kernel), algorithmic bytes per
element-step = 480 FP64 values read/written (3840 B) + 112 B geometry + 16 B
level indices = 3968 B, divided by the sweep's device time (HIP events on
the solver stream), against 8 TB/s.
cpu_baseline: the oracle's restatement of the reference CPU algorithm
(assembled operators, sparse LU per velocity) on a bounded sample of the same
workload: the first few steps at 32^3, assembly excluded.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dg-cfl-less-maxwell-solver_amd"))

import numpy as np  # noqa: E402

import gdon3d as G  # noqa: E402

METRIC = "DoF-updates/s at fixed L2 error, P2 Maxwell RK3 & relaxation; HBM GB/s"
HBM_PEAK_GBS = 8000.0
SWEEP_BYTES_PER_ELEM = 480 * 8 + 112 + 16
# HBM bytes actually moved per launch, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE passes calibrated against a 1 GiB copy (FETCH_SIZE reads half the
# bytes, WRITE_SIZE all of them); per element of the workloads measured there
# (profiles/hbm_traffic_sweep_relax_rk3.txt): the sweep of one step at 32^3
# D3Q4 and one RK3 stage at 64^3 jittered.
SWEEP_TRAFFIC_BYTES_PER_ELEM = 6.245e9 / 196608
RK3_TRAFFIC_BYTES_PER_ELEM_STAGE = (2 * 2634964.1 + 1474560.0) * 1024 / 1572864


def cpu_baseline(n, steps):
    """Oracle (C99 + OpenMP restatement of the reference) on host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import tempfile

    import oracle as O

    path = os.path.join(tempfile.mkdtemp(), f"kuhn{n}.msh")
    G.load_library().mesh_kuhn_cube_write_msh22(path.encode(), n, 0.0, 0)
    om = O.OMesh(path)
    par = np.zeros(16)
    par[0:3] = np.pi
    par[3], par[4], par[6] = 1.0, -1.0, np.pi * np.sqrt(3.0)
    # the reference parallelises the transport over the 4 velocities only
    threads = min(4, os.cpu_count() or 1)
    s = O.OSim(om, 1, O.CAVITY_MODE, par, kin=O.D3Q4, threads=threads)
    s.set_time(2.0, 0.0, 5.0)
    s.solve(steps)
    dof = 24 * 10 * om.nelem
    out = dict(value=dof * steps / s.t_run, unit="DoF-updates/s", cores=threads,
               kind="port",
               sample=f"oracle relaxation D3Q4 FP64 on Kuhn {n}^3x6, first {steps} "
                      f"steps timed ({s.t_run:.2f} s; assembly {s.t_asm:.2f} s excluded); "
                      f"transport parallel over the 4 velocities, relaxation serial, "
                      f"as the reference")
    s.free()
    om.free()
    return out


RK3_BYTES_PER_ELEM_STAGE = 4 * 60 * 8 + 112  # read w, k2; write k2, w; geometry


def bench_rk3(args):
    """RK3, plane wave k = 2 pi (1,1,0)/sqrt 2 with E along z, Kuhn 64^3 x 6
    with +-10% interior jitter (seed 17), CFL 0.5; --precision fp32 is the
    precision form (binary32 fields and flux/volume arithmetic)."""
    n = args.n if args.n != 32 else 64
    k = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)
    G.set_plane_wave(k, (0.0, 0.0, 1.0))
    mesh = G.Mesh(kuhn=(n, 0.1, 17))
    sim = G.Simulation(mesh, G.Model("plane_wave"))
    sim.set_time(1.0, 0.0, 0.5)
    nelem = mesh.nelem
    dof = 6 * 10 * nelem
    st = G.Stepper(sim, precision=args.precision)
    st.advance(args.warmup)
    st.sync()
    t0 = time.perf_counter()
    st.advance(args.steps)
    st.sync()
    elapsed = time.perf_counter() - t0
    err = st.error_l2()
    st.profile(True)
    st.advance(args.profile_steps)
    prof = st.profile(False)
    stage_s = prof[0] / (3 * args.profile_steps)
    s_val = 4 if args.precision == "fp32" else 8
    bytes_elem = 4 * 60 * s_val + 112
    achieved = bytes_elem * nelem / stage_s / 1e9
    print(f"[bench rk3] n={n} elements={nelem} dt={sim.s.dt:.4e} {args.steps} steps "
          f"{elapsed:.3f}s stage kernel {stage_s*1e3:.3f} ms ghosts "
          f"{prof[2]/args.profile_steps*1e3:.3f} ms/step L2 {err:.3e}",
          file=sys.stderr, flush=True)
    st.free()
    sim.free()
    line = {
        "metric": METRIC, "value": dof * args.steps / elapsed, "unit": "DoF-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"low-storage RK3, {args.precision.upper()}, plane wave k=2pi(1,1,0)/sqrt2 E||z, "
                               f"Kuhn {n}^3x6 = {nelem} P2 tets, jitter 10% seed 17, CFL 0.5",
                   "elements": nelem, "dof": dof, 
                   "l2_error": err},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (RK3_TRAFFIC_BYTES_PER_ELEM_STAGE * nelem
                                 if args.precision == "fp64" and n == 64 else None),
                     "kernel": "k_rk3_soa", "bytes_per_element_stage": bytes_elem,
                     "avg_launch_us": stage_s * 1e6},
    }
    print(json.dumps(line), flush=True)


def copy_calibration(args):
    """A plain device copy of known size (1 GiB read + 1 GiB written per
    launch, kernel k_copy_calib), run under the same rocprofv3 --pmc passes
    as the solver kernels: FETCH_SIZE / WRITE_SIZE of this copy over 1 GiB is
    the factor by which the counters over- or under-count HBM bytes."""
    lib = G.load_library()
    ms = lib.gdh_copy_calibration(1 << 30, args.steps)
    print(f"[copy] 1 GiB copy: {ms:.3f} ms, {2 * (1 << 30) / ms / 1e6:.1f} GB/s",
          file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=316)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="fp32: the precision form (binary32 fields)")
    ap.add_argument("--jitter", type=float, default=0.0,
                    help="relax: interior vertex jitter (fraction of h, seed 17)")
    ap.add_argument("--exact", choices=["cavity", "plane"], default="cavity",
                    help="relax: PEC cavity mode (1,1,1) to t=2, or the oblique "
                         "plane wave k=2pi(1,1,0)/sqrt2 E||z to t=1")
    ap.add_argument("--scheme", choices=["relax", "rk3", "copy"], default="relax",
                    help="relax: the headline workload; rk3: low-storage RK3 "
                         "on the jittered 64^3 plane-wave mesh (secondary); copy: "
                         "1 GiB device copy that calibrates the HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the design for sharding is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    lib = G.load_library()
    if lib.gdh_device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    if args.scheme == "rk3":
        return bench_rk3(args)
    if args.scheme == "copy":
        return copy_calibration(args)
    n = args.n
    if args.exact == "cavity":
        G.set_cavity_mode(1, 1, 1, (1.0, -1.0, 0.0))
        mesh = G.Mesh(kuhn=(n, args.jitter, 17))
        model = G.Model("cavity_mode", kinetic="d3q4")
        tend = 2.0
        what = "PEC cube mode (1,1,1)"
    else:
        kv = 2 * np.pi * np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)
        G.set_plane_wave(kv, (0.0, 0.0, 1.0))
        mesh = G.Mesh(kuhn=(n, args.jitter, 17))
        model = G.Model("plane_wave", kinetic="d3q4")
        tend = 1.0
        what = "plane wave k=2pi(1,1,0)/sqrt2 E||z"
    sim = G.Simulation(mesh, model)
    sim.set_time(tend, 0.0, 5.0)
    nelem = mesh.nelem
    dof = 24 * 10 * nelem

    t0 = time.perf_counter()
    st = G.Stepper(sim, precision=args.precision)
    st.sync()
    t_setup = time.perf_counter() - t0
    sim_itermax = sim.s.itermax
    default_workload = (n == 32 and not args.jitter and args.exact == "cavity"
                        and args.precision == "fp64")
    print(f"[bench] n={n} elements={nelem} itermax={sim.s.itermax} dt={sim.s.dt:.6e} "
          f"levels={st.levels} setup={t_setup:.2f}s", file=sys.stderr, flush=True)

    st.advance(args.warmup)
    st.sync()
    t0 = time.perf_counter()
    st.advance(args.steps)
    st.sync()
    elapsed = time.perf_counter() - t0
    err = st.error_l2()
    t_final = st.t
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f}s, t={t_final:.5f}, "
          f"L2 error {err:.4e}", file=sys.stderr, flush=True)

    # per-kernel device time (HIP events on the solver stream)
    st.profile(True)
    st.advance(args.profile_steps)
    prof = st.profile(False)
    sweep_s = prof[0] / args.profile_steps
    relax_s = prof[1] / args.profile_steps
    launches = prof[3] / args.profile_steps
    s_val = 4 if args.precision == "fp32" else 8
    sweep_bytes = 480 * s_val + 112 + 16
    achieved = sweep_bytes * nelem / sweep_s / 1e9
    print(f"[bench] per step: sweep {sweep_s*1e3:.3f} ms ({launches:.0f} launches), "
          f"relax {relax_s*1e3:.3f} ms, ghosts {prof[2]/args.profile_steps*1e3:.3f} ms",
          file=sys.stderr, flush=True)
    st.free()
    sim.free()

    value = dof * args.steps / elapsed
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "DoF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"CFL-less relaxation, D3Q4, {args.precision.upper()}, {what}, "
                        f"Kuhn {n}^3x6 = {nelem} P2 tets"
                        f"{f', jitter {args.jitter:g} seed 17' if args.jitter else ''}, "
                        f"cfl 5, t_end {tend:g}",
            "elements": nelem,
            "dof": dof,
            "itermax": sim_itermax,
            "t_final": t_final,
            "l2_error": err,
            "hbm_gb_s_sweep": achieved,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (SWEEP_TRAFFIC_BYTES_PER_ELEM * nelem if default_workload else None),
            "traffic_gb_s": (SWEEP_TRAFFIC_BYTES_PER_ELEM * nelem / sweep_s / 1e9
                             if default_workload else None),
            "kernel": "k_sweep_persistent (one launch = the whole sweep of one step)",
            "bytes_per_element_step": sweep_bytes,
            "launches_per_step": launches,
            "avg_launch_us": sweep_s / max(launches, 1) * 1e6,
        },
    }
    if not args.no_cpu_baseline and default_workload:
        line["cpu_baseline"] = cpu_baseline(n, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
