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
roofline: the kernels of one step, the transport sweep (one launch) and,
when it runs as a kernel of its own, the relaxation pass. Algorithmic bytes
per DoF-update (s = 8 FP64, 4 FP32): transport reads and writes the 24
micro-fields (2 s), the relaxation reads and writes them (2 s), plus 112 B
geometry + 16 B level indices per element (128 / 240 per DoF): 4 s + 0.53 =
32.5 B (FP64), 16.5 B (FP32), (960 s + 128) B per element and step.
achieved = those bytes / the mean device time of the step's kernels (HIP
events on the solver stream: the sweep with the relaxation fused into it, or
the sweep plus the pass), against 8 TB/s; achieved_step = the same bytes x
DoF-updates/s over the whole step (ghosts and launch gaps included);
achieved_transport_only = the transport half over the sweep kernel alone.
cpu_baseline: the oracle's restatement of the reference CPU algorithm
(assembled operators, sparse LU per velocity) on a bounded sample of the same
workload: the first steps at 32^3, assembly excluded.
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
from workloads import batch_instances, learned_eps  # noqa: E402

METRIC = "DoF-updates/s at fixed L2 error, P2 Maxwell RK3 & relaxation; HBM GB/s"
HBM_PEAK_GBS = 8000.0
SWEEP_BYTES_PER_ELEM = 480 * 8 + 112 + 16
# HBM bytes actually moved per launch, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE passes calibrated against a 1 GiB copy (FETCH_SIZE reads half the
# bytes, WRITE_SIZE all of them); per element of the workloads measured there
# (profiles/hbm_traffic_sweep_relax_rk3.txt): the sweep of one step at 32^3
# D3Q4 and one RK3 stage at 64^3 jittered.
SWEEP_TRAFFIC_BYTES_PER_ELEM = {  # (sweep form, relaxation fused into the sweep)
    ("item", 1): (1714364.3 / 0.5 + 737613.4) * 1024 / 196608,
    ("item", 0): (1031197.7 / 0.5 + 737763.0) * 1024 / 196608,
    ("cluster", 0): (915513.6 / 0.5 + 646294.6) * 1024 / 196608,
}
# the relaxation pass of one step (k_relax_pair<4, double>) at 32^3 D3Q4
RELAX_TRAFFIC_BYTES_PER_ELEM = (184362.8 / 0.5 + 368892.1) * 1024 / 196608
SWEEP_KERNEL = {"level": "k_sweep_level", "item": "k_sweep_persistent",
                "cluster": "k_sweep_cluster"}
# gdh_relax_fused: where the relaxation of a step is applied
RELAX_FORMS = {0: "in a pass of its own after the sweep",
               1: "fused into the next sweep's reads",
               2: "by each element's last item inside the sweep"}
RK3_TRAFFIC_BYTES_PER_ELEM_STAGE = {  # (precision, stage kernel)
    ("fp64", "k_rk3_pair"): (1112042.4 / 0.5 + 1474595.8) * 1024 / 1572864,
    ("fp32", "k_rk3_tile"): (668114.9 / 0.5 + 737280.2) * 1024 / 1572864,
}


def host_cpu():
    """(logical CPUs of this host, CPU model name from /proc/cpuinfo)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, model


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
    nproc, model = host_cpu()
    out = dict(value=dof * steps / s.t_run, unit="DoF-updates/s", cores=threads,
               kind="port", host_nproc=nproc, host_cpu_model=model,
               sample=f"oracle relaxation D3Q4 FP64 on Kuhn {n}^3x6, first {steps} "
                      f"steps timed ({s.t_run:.2f} s; assembly {s.t_asm:.2f} s excluded); "
                      f"transport parallel over the 4 velocities, relaxation serial, "
                      f"as the reference; {threads} of {nproc} host threads ({model}); "
                      f"the KLU variant is not timed: SuiteSparse is not installed")
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
    # the stage kernel the library runs: two lanes per element (binary64
    # default), one (binary32 default), or GDN_RK3's choice
    kernel = {"soa": "k_rk3_soa", "pair": "k_rk3_pair", "tile": "k_rk3_tile"}.get(
        os.environ.get("GDN_RK3", ""), "k_rk3_tile" if args.precision == "fp32" else "k_rk3_pair")
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
                     "traffic": (RK3_TRAFFIC_BYTES_PER_ELEM_STAGE[(args.precision, kernel)] * nelem
                                 if n == 64 and (args.precision, kernel)
                                 in RK3_TRAFFIC_BYTES_PER_ELEM_STAGE else None),
                     "kernel": kernel,
                     "bytes_per_element_stage": bytes_elem,
                     "avg_launch_us": stage_s * 1e6,
                     "copy_peak_measured": copy_peak_gbs()},
    }
    print(json.dumps(line), flush=True)


def copy_peak_gbs(reps=20):
    """Measured device copy rate in GB/s (1 GiB read + 1 GiB written per
    launch, kernel k_copy_calib), taken after the timed region: the practical
    HBM ceiling reported beside the 8 TB/s specification."""
    ms = G.load_library().gdh_copy_calibration(1 << 30, reps)
    return 2 * (1 << 30) / ms / 1e6 if ms > 0 else None


def copy_calibration(args):
    """A plain device copy of known size (1 GiB read + 1 GiB written per
    launch, kernel k_copy_calib), run under the same rocprofv3 --pmc passes
    as the solver kernels: FETCH_SIZE / WRITE_SIZE of this copy over 1 GiB is
    the factor by which the counters over- or under-count HBM bytes."""
    lib = G.load_library()
    ms = lib.gdh_copy_calibration(1 << 30, args.steps)
    print(f"[copy] 1 GiB copy: {ms:.3f} ms, {2 * (1 << 30) / ms / 1e6:.1f} GB/s",
          file=sys.stderr, flush=True)


def bench_batch(args):
    """Batched form: 256 instances (workloads.batch_instances), FP32 fields
    by default, all instances (and velocities) per launch, captured-graph
    steps, autotuned tiles, eps per 4^3 block of cubes. --batch-kind rk3
    (CFL 0.5) or relax (D3Q4, cfl 5)."""
    kin = args.batch_kind == "relax"
    inst = batch_instances(args.instances)
    t0 = time.perf_counter()
    sims, pars, eps, dof = [], [], [], 0
    for n, par, e in inst:
        mesh = G.Mesh(kuhn=(n, 0.0, 0))
        sim = G.Simulation(mesh, G.Model("plane_wave", kinetic="d3q4" if kin else None))
        sim.set_time(1.0, 0.0, 5.0 if kin else 0.5)
        sims.append(sim)
        pars.append(par)
        eps.append(e)
        dof += (24 if kin else 6) * 10 * 6 * n ** 3
    nelem = sum(6 * n ** 3 for n, _, _ in inst)
    seq = None
    if args.batch_sequential:
        # the same instances one stepper each, advanced one after the
        # other: what running them without the batch costs
        steppers = []
        for sm, par, e in zip(sims, pars, eps):
            G.set_plane_wave(par[0:3], par[3:6])
            steppers.append(G.Stepper(sm, precision=args.precision, eps=e))
        for st in steppers:
            st.advance(2)
        for st in steppers:
            st.sync()
        ts = time.perf_counter()
        for st in steppers:
            st.advance(args.steps)
        for st in steppers:
            st.sync()
        seq = dof * args.steps / (time.perf_counter() - ts)
        for st in steppers:
            st.free()
    b = G.Batch(sims, np.array(pars), eps=np.concatenate(eps), precision=args.precision)
    b.sync()
    t_setup = time.perf_counter() - t0
    tiles = b.tiles()
    b.advance(args.warmup)
    b.sync()
    t0 = time.perf_counter()
    b.advance(args.steps)
    b.sync()
    elapsed = time.perf_counter() - t0
    err = b.errors()
    b.free()
    for sm in sims:
        sm.free()
    ns = np.array([n for n, _, _ in inst])
    uni = ns <= 4  # one 4^3 block of eps: the exact solution is known
    value = dof * args.steps / elapsed
    s_val = 4 if args.precision == "fp32" else 8
    b_alg = (4 * s_val + 128 / 240) if kin else 3 * (240 * s_val + 112) / 60
    achieved = value * b_alg / 1e9
    print(f"[bench batch] {args.instances} instances, n in [{ns.min()}, {ns.max()}], "
          f"{nelem} tets, setup {t_setup:.1f}s, tiles {tiles}, {args.steps} steps "
          f"{elapsed:.3f}s", file=sys.stderr, flush=True)
    line = {
        "metric": METRIC, "value": value, "unit": "DoF-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"batched form: {args.instances} Kuhn cubes, n log-uniform "
                               f"in [4,24] (seed 5), random plane waves |k|=2pi m, "
                               + ("D3Q4 relaxation cfl 5" if kin else "RK3 CFL 0.5")
                               + ", eps in [1,4] per 4^3 block"
                               + f", {args.precision.upper()}, captured-graph steps",
                   "instances": args.instances, "elements": nelem, "dof": dof,
                   "tiles_rk3_relax": list(tiles),
                   "sequential_steppers_dof_updates_per_s": seq,
                   "uniform_eps_instances": int(uni.sum()),
                   "l2_error_median": float(np.median(err[uni])) if uni.any() else None,
                   "l2_error_max": float(err[uni].max()) if uni.any() else None,
                   "l2_note": "L2 error against the exact plane wave of speed 1/sqrt(eps), "
                              "over the instances of one eps block (n <= 4); the others "
                              "have eps jumping between blocks and no closed-form solution"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "whole step (all kernels, graph replay)",
                     "bytes_per_dof_update": b_alg,
                     "copy_peak_measured": copy_peak_gbs()},
    }
    print(json.dumps(line), flush=True)


def sweep_ms(st, steps):
    """mean device time of the transport sweep per step over `steps` steps
    (HIP events on the solver stream)"""
    st.profile(True)
    st.advance(steps)
    prof = st.profile(False)
    return prof[0] / steps * 1e3


def bench_learned(args):
    """Learned-preconditioner form: relaxation (D3Q4) on Kuhn 48^3 x 6 with
    +-20% jitter (seed 23), eps uniform in [1, 4] per element (seed 23), a
    Gaussian pulse E_z (sigma 0.08, centred), cfl 10 = 20x the explicit CFL.
    The block network is trained on the GPU on 10^5 items of this mesh
    (features and exact block inverses). Reported: Richardson block solves
    (6 random sides per item on 10^5 other items, to |r - Du| <= 1e-8 |r|,
    exact fallback) with the network and with the identity; then the
    relaxation steps twice from the same initial state, with the stored
    exact block inverses and with the network's inverses refined in the
    sweep (gdh_learned_enable): DoF-updates/s of both, the items on the
    stored inverse (fallbacks) and the relative L2 difference of the two
    solutions. value = the learned form's rate."""
    n = args.n if args.n != 32 else 48
    G.set_gaussian_pulse((0.5, 0.5, 0.5), 0.08, 1.0)
    mesh = G.Mesh(kuhn=(n, 0.2, 23))
    nelem = mesh.nelem
    eps = learned_eps(nelem, 23)
    sim = G.Simulation(mesh, G.Model("gaussian_pulse", kinetic="d3q4"))
    sim.set_time(1.0, 0.0, 10.0)
    t0 = time.perf_counter()
    st = G.Stepper(sim, precision=args.precision, eps=eps)
    st.sync()
    t_setup = time.perf_counter() - t0
    lagged = st.transport_stats()[0]
    nitems = 4 * nelem
    rng = np.random.default_rng(23)
    perm = rng.permutation(nitems)
    train_idx, test_idx = perm[:100000], perm[100000:200000]
    x, y = G.learned_data(st, train_idx)
    net = G.BlockNet(seed=5)
    net.set_normalization(G.normalization(x, y))
    t0 = time.perf_counter()
    hist = net.train(x, y, steps=args.train_steps, batch=1024, lr=2e-3)
    t_train = time.perf_counter() - t0
    stats = G.learned_solve_stats(st, net, test_idx)
    ident = G.learned_solve_stats(st, None, test_idx)
    gm = G.learned_solve_stats(st, net, test_idx, method="gmres")
    gm_id = G.learned_solve_stats(st, None, test_idx, method="gmres")
    dof = 24 * 10 * nelem
    runs = {}
    for mode in ("exact", "learned"):
        lstats = G.learned_enable(st, net if mode == "learned" else None)
        form = (G.sweep_info(st)["form"], G.relax_fused(st))
        G.load_library().dg_stepper_init(st.h)
        st.advance(args.warmup)
        st.sync()
        t0 = time.perf_counter()
        st.advance(args.steps)
        st.sync()
        elapsed = time.perf_counter() - t0
        wn = st.download()
        ms = sweep_ms(st, args.profile_steps)
        runs[mode] = dict(value=dof * args.steps / elapsed, elapsed=elapsed, wn=wn,
                          sweep_ms=ms, stats=lstats, form=form)
    diff = float(np.linalg.norm(runs["learned"]["wn"] - runs["exact"]["wn"]) /
                 np.linalg.norm(runs["exact"]["wn"]))
    t_end = st.t
    energy = st.energy()
    net.free()
    st.free()
    sim.free()
    lr, ex = runs["learned"], runs["exact"]
    print(f"[bench learned] n={n} {nelem} tets, setup {t_setup:.1f}s, upwind cycles: "
          f"{lagged} lagged faces; training {args.train_steps} steps {t_train:.2f}s loss "
          f"{hist[:5].mean():.3e} -> {hist[-20:].mean():.3e}; Richardson block solves: mean "
          f"{stats[0]:.2f} it, max {stats[1]:.0f}, fallback {stats[2]:.4f}, max rel L2 "
          f"{stats[3]:.2e} (identity preconditioner: fallback {ident[2]:.3f}); sweep "
          f"exact {ex['sweep_ms']:.3f} ms ({ex['form']}), learned {lr['sweep_ms']:.3f} ms "
          f"({lr['form']}), items on the network / stored inverse {lr['stats'][:2]}, "
          f"max deviation {lr['stats'][2]:.2e}; learned vs exact solution rel L2 {diff:.2e}",
          file=sys.stderr, flush=True)
    line = {
        "metric": METRIC, "value": lr["value"], "unit": "DoF-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": lr["elapsed"] / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"learned-preconditioner form: D3Q4 relaxation, Kuhn {n}^3x6 "
                               f"= {nelem} tets, jitter 20% seed 23, eps in [1,4] per element "
                               f"(seed 23), Gaussian pulse E_z sigma 0.08, cfl 10 (20x "
                               f"explicit); block inverses from the network refined in the "
                               f"sweep, stored exact inverse as fallback",
                   "elements": nelem, "lagged_cycle_faces": lagged,
                   "train_steps": args.train_steps, "train_s": t_train,
                   "loss_first": float(hist[:5].mean()), "loss_last": float(hist[-20:].mean()),
                   "richardson_mean_iterations": float(stats[0]),
                   "richardson_max_iterations": float(stats[1]),
                   "richardson_fallback_rate": float(stats[2]),
                   "richardson_max_rel_l2_vs_exact": float(stats[3]),
                   "identity_preconditioner_fallback_rate": float(ident[2]),
                   "gmres_mean_products": float(gm[0]),
                   "gmres_max_products": float(gm[1]),
                   "gmres_fallback_rate": float(gm[2]),
                   "gmres_identity_mean_products": float(gm_id[0]),
                   "items_on_network_inverse": lr["stats"][0],
                   "items_on_stored_inverse": lr["stats"][1],
                   "refined_inverse_max_rel_deviation": lr["stats"][2],
                   "exact_inverse_dof_updates_per_s": ex["value"],
                   "gain_over_exact_inverse": lr["value"] / ex["value"],
                   "sweep_ms_exact": ex["sweep_ms"], "sweep_ms_learned": lr["sweep_ms"],
                   "sweep_form_exact": list(ex["form"]), "sweep_form_learned": list(lr["form"]),
                   "learned_vs_exact_solution_rel_l2": diff,
                   "t_final": t_end, "energy_final": energy},
        "roofline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=316)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--energy-every", type=int, default=10,
                    help="relax: energy history sampled every k steps over a second, "
                         "untimed run to t_end (0: none)")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="fp32: the precision form (binary32 fields)")
    ap.add_argument("--jitter", type=float, default=0.0,
                    help="relax: interior vertex jitter (fraction of h, seed 17)")
    ap.add_argument("--exact", choices=["cavity", "plane"], default="cavity",
                    help="relax: PEC cavity mode (1,1,1) to t=2, or the oblique "
                         "plane wave k=2pi(1,1,0)/sqrt2 E||z to t=1")
    ap.add_argument("--instances", type=int, default=256, help="batch: instances")
    ap.add_argument("--train-steps", type=int, default=2000, help="learned: Adam steps")
    ap.add_argument("--batch-kind", choices=["rk3", "relax"], default="rk3")
    ap.add_argument("--learned-inverses", action="store_true",
                    help="relax: the sweep's block inverses from a network trained on 1e5 "
                         "items of the mesh, refined in the sweep (gdh_learned_enable)")
    ap.add_argument("--batch-sequential", action="store_true",
                    help="batch: also time the instances one stepper each")
    ap.add_argument("--scheme", choices=["relax", "rk3", "copy", "batch", "learned"],
                    default="relax",
                    help="relax: the headline workload; rk3: low-storage RK3 "
                         "on the jittered 64^3 plane-wave mesh (secondary); copy: "
                         "1 GiB device copy that calibrates the HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the solver has no multi-GPU path "
              "(sharding by velocity or by element partition is not implemented)", file=sys.stderr)
        sys.exit(2)

    lib = G.load_library()
    if lib.gdh_device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    if args.scheme == "rk3":
        return bench_rk3(args)
    if args.scheme == "copy":
        return copy_calibration(args)
    if args.scheme == "batch":
        return bench_batch(args)
    if args.scheme == "learned":
        return bench_learned(args)
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
    learned = None
    if args.learned_inverses:
        idx = np.random.default_rng(23).permutation(4 * nelem)[:100000]
        x, y = G.learned_data(st, idx)
        net = G.BlockNet(seed=5)
        net.set_normalization(G.normalization(x, y))
        net.train(x, y, steps=args.train_steps, batch=1024, lr=2e-3)
        learned = G.learned_enable(st, net)
        net.free()
        G.load_library().dg_stepper_init(st.h)
    relax_fused = G.relax_fused(st)
    sweep = G.sweep_info(st)
    traffic = SWEEP_TRAFFIC_BYTES_PER_ELEM.get((sweep["form"], relax_fused))
    sim_itermax = sim.s.itermax
    default_workload = (n == 32 and not args.jitter and args.exact == "cavity"
                        and args.precision == "fp64" and not args.learned_inverses)
    # measured HBM bytes of one step's kernels (the headline workload only)
    step_traffic = None
    if default_workload and traffic:
        step_traffic = traffic + (0.0 if relax_fused else RELAX_TRAFFIC_BYTES_PER_ELEM)
    print(f"[bench] n={n} elements={nelem} itermax={sim.s.itermax} dt={sim.s.dt:.6e} "
          f"levels={st.levels} sweep={sweep} "
          f"relax_fused={relax_fused} setup={t_setup:.2f}s",
          file=sys.stderr, flush=True)

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
    # algorithmic bytes of one step per element: the transport and the
    # relaxation each read and write the 24 micro-fields, plus geometry and
    # schedule (32.5 B per DoF-update in binary64, 16.5 B in binary32),
    # over the device time of that step's kernels: the sweep, and the
    # relaxation pass when it runs as a kernel of its own (the same figure
    # whichever form the set-up chose)
    transport_bytes = 480 * s_val + 112 + 16
    step_bytes = 960 * s_val + 128
    launch_s = sweep_s / max(launches, 1)
    kernels_s = sweep_s + relax_s
    achieved = step_bytes * nelem / kernels_s / 1e9
    print(f"[bench] per step: sweep {sweep_s*1e3:.3f} ms ({launches:.0f} launches), "
          f"relax {relax_s*1e3:.3f} ms, ghosts {prof[2]/args.profile_steps*1e3:.3f} ms",
          file=sys.stderr, flush=True)
    # energy history, untimed: the whole run from t = 0 to t_end again, the
    # discrete energy 1/2 w^T M w (binary64 device reduction) after every step
    hist_t, hist_e = [], []
    if args.energy_every > 0:
        G.load_library().dg_stepper_init(st.h)
        hist_t.append(0.0)
        hist_e.append(st.energy())
        for k in range(1, sim_itermax + 1):
            st.advance(1)
            if k % args.energy_every == 0 or k == sim_itermax:
                hist_t.append(st.t)
                hist_e.append(st.energy())
        print(f"[bench] energy history: {len(hist_e)} samples, E(0) = {hist_e[0]:.10e}, "
              f"E(t_end = {hist_t[-1]:.5f}) = {hist_e[-1]:.10e} (ratio {hist_e[-1]/hist_e[0]:.8f})",
              file=sys.stderr, flush=True)
    st.free()
    sim.free()

    value = dof * args.steps / elapsed
    b_alg = 4 * s_val + 128 / 240
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
            "hbm_gb_s": achieved,
            "sweep_ms": sweep_s * 1e3,
            "relaxation_pass_ms": relax_s * 1e3,
            "relaxation_fused_into_sweep": relax_fused == 1,
            "relaxation_form": RELAX_FORMS.get(relax_fused, str(relax_fused)),
            "sweep_form": sweep["form"],
            "sweep_clusters": sweep["clusters"],
            "sweep_resident_workgroups": sweep["grid"],
            "setup_s": t_setup,
            "learned_inverses": (None if learned is None else
                                 {"items_on_network": learned[0], "items_on_stored": learned[1],
                                  "max_rel_deviation": learned[2]}),
            "energy_history": ({"t": hist_t, "energy": hist_e,
                                "ratio_end": hist_e[-1] / hist_e[0],
                                "note": "1/2 w^T M w, binary64 device reduction, second "
                                        "untimed run from t = 0"} if hist_e else None),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": step_traffic * nelem if step_traffic else None,
            "traffic_gb_s": (step_traffic * nelem / kernels_s / 1e9 if step_traffic else None),
            "kernel": (f"{SWEEP_KERNEL[sweep['form']]} (one launch = the whole sweep of one step, "
                       f"relaxation {RELAX_FORMS[relax_fused]})" if relax_fused else
                       f"{SWEEP_KERNEL[sweep['form']]} + k_relax (the sweep and the relaxation "
                       f"pass of one step)"),
            "bytes_per_element_step": step_bytes,
            "bytes_per_dof_update": step_bytes / 240,
            "launches_per_step": launches,
            "avg_launch_us": launch_s * 1e6,
            "relaxation_pass_us": relax_s * 1e6,
            "step_kernels_us": kernels_s * 1e6,
            "achieved_step": value * b_alg / 1e9,
            "frac_step": value * b_alg / 1e9 / HBM_PEAK_GBS,
            "bytes_per_dof_update_step": b_alg,
            "achieved_transport_only": transport_bytes * nelem / sweep_s / 1e9,
            "copy_peak_measured": copy_peak_gbs(),
        },
    }
    if not args.no_cpu_baseline and default_workload:
        line["cpu_baseline"] = cpu_baseline(n, args.cpu_steps)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
