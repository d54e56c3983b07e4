#!/usr/bin/env python3
"""Benchmark of the coupled platelet step on one MI355X.

Workload: single platelet in shear, 256x128x128 D3Q19 Couette flow (wall
velocity 0.05, tau 0.7), one 2048-element SEM platelet with 512 receptors over
4096 wall ligands, Bell bonds, FP32 storage.  A step is one full coupled step:
cell lists, membrane sums, bonds, forces, IB spreading, fused LB
collide-stream, IB interpolation and marker update.  Inputs are resident in
HBM before the timed region; the timed region is K hipGraph step launches
between two device synchronisations.

Prints ONE JSON line.  value = coupled-step MLUPS (lattice cells x steps / s).
roofline: the fused LB kernel (the dominant kernel), algorithmic bytes
152 B/cell (FP32 storage: 19 x 4 B read + 19 x 4 B write) over its mean
duration, measured with HIP events on the solver's stream around each of K
back-to-back launches of the same kernel on the same lattice.

Other modes (for rocprofv3 runs):
  --lb-only f32|f16|bf16   K fused LB steps only, no markers
  --copy-probe             the library's float4 streaming copy (2 x 318 MB)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {0: 152, 1: 76, 2: 76}
STORES = {"f32": 0, "f16": 1, "bf16": 2}
METRIC = "D3Q19 MLUPS, % HBM peak; coupled steps/s FP16-store vs FP32 at stated tol"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "lb_fused_shear256_traffic.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_solver(pfg, scene, store, fluid_only=False):
    p = scene.params
    p.store = store
    s = pfg.Solver(p)
    s.load(scene)
    s.set_fluid_only(fluid_only)
    return s


TUNING = {}
NO_TUNE = False  # --no-tune: keep the automatic launch shapes (profiling runs)


def timed_run(pfg, scene, store, steps, warmup, fluid_only=False, tune=None):
    """Warm-up steps, then `steps` timed steps.  With `tune` (a key), the
    launch shapes are picked first by the library's autotuner, outside the
    timed region, and its choice is kept in TUNING[tune]."""
    s = make_solver(pfg, scene, store, fluid_only)
    if tune is not None and not NO_TUNE:
        TUNING[tune] = s.autotune(32)
        log(f"autotune {tune}: {TUNING[tune]}")
    s.step(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    dt = time.perf_counter() - t0
    return s, dt


def lb_kernel_us(pfg, scene, store, launches):
    """Mean LB-kernel duration: HIP events on the solver's stream around
    `launches` back-to-back launches of the LB kernel."""
    s = make_solver(pfg, scene, store, fluid_only=True)
    s.step(10)
    ms = s.time_lb(launches)
    s.close()
    return ms * 1e3


def read_traffic():
    """HBM bytes per LB launch from the committed rocprofv3 PMC summary."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        key, _, val = line.partition("=")
        if key.strip() == "traffic_bytes_per_launch":
            try:
                return float(val.split()[0])
            except (IndexError, ValueError):
                return None
    return None


def oracle_threads():
    """Host threads for the FP64 oracle: the process's CPU share, at most 16
    (OMP_NUM_THREADS when it is set)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def load_oracle():
    """The FP64 oracle, used only in the CPU leg below: timed as the CPU
    baseline and, as the checker, for the accuracy figures of the GPU runs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = pyoracle.set_threads(oracle_threads())
    return pyoracle, threads


def time_oracle(o, cells, budget_s=10.0, min_steps=2, max_steps=200):
    """Coupled oracle steps until about `budget_s` of wall time: MLUPS, steps,
    seconds."""
    t0 = time.perf_counter()
    steps = 0
    while steps < min_steps or (time.perf_counter() - t0 < budget_s and steps < max_steps):
        o.step(1)
        steps += 1
    dt = time.perf_counter() - t0
    return cells * steps / dt / 1e6, steps, dt


def cpu_leg_shear(pfg):
    """CPU leg of the headline workload.  Baseline: the FP64 oracle (the same
    algorithm, OpenMP over cells) on a bounded sample of the same workload,
    about 10 s of coupled steps of the 256x128x128 shear scene.  Checker: the
    FP16-store (and BF16-store) velocity error against the oracle on a
    64x32x32 body-force channel, tau 0.8, F = 1e-6, 300 steps."""
    import numpy as np
    po, threads = load_oracle()
    scene = pfg.Scene(pfg.SCENE_SHEAR)
    o = po.Oracle(scene.params, scene=scene)
    mlups, steps, dt = time_oracle(o, scene.N)
    o.close()
    cpu = {"value": round(mlups, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
           "sample": f"{steps} coupled steps of the 256x128x128 shear workload, FP64 oracle, "
                     f"{threads} OpenMP threads ({dt:.1f} s)"}
    p = pfg.default_params(nx=64, ny=32, nz=32, tau=0.8, store=pfg.STORE_F16)
    p.wall[1] = p.wall[2] = pfg.WALL_BOUNCE
    p.body_force[0] = 1e-6
    L = po.lattice(p)
    f = po.init_eq(L)
    for _ in range(300):
        f = po.lb_step(L, f)
    _, ox, oy, oz = po.fields(L, f)
    den = np.sqrt((ox ** 2 + oy ** 2 + oz ** 2).sum())
    err = {}
    for store in (pfg.STORE_F16, pfg.STORE_BF16):
        p.store = store
        with pfg.Solver(p) as s:
            s.set_fluid_only(True)
            s.step(300)
            _, ux, uy, uz = s.fields()
        num = np.sqrt(((ux - ox) ** 2 + (uy - oy) ** 2 + (uz - oz) ** 2).sum())
        err[store] = float(num / den)
    return cpu, err[pfg.STORE_F16], err[pfg.STORE_BF16]


def subbox_scene(pfg, n=128):
    """The large configuration's physics on an n^3 sub-box with the same
    per-cell parameters: tau 0.7, Couette with the shear rate of a 0.05 wall
    velocity over 512 cells, FP16 deviation storage, the same densities of
    platelets (2000 x 1000 elements per 1024x512x512), fibrin (80k nodes,
    120k links) and wall ligands (65536 per 1024x512)."""
    import numpy as np
    frac = n ** 3 / (1024 * 512 * 512)
    sc = pfg.Scene(pfg.SCENE_LARGE, nx=n, ny=n, nz=n,
                   n_platelets=max(1, round(2000 * frac)),
                   fibrin_nodes=max(8, round(80000 * frac)),
                   fibrin_links=max(8, round(120000 * frac)),
                   ligands=max(64, round(65536 * n * n / (1024 * 512))))
    uw = 0.05 * n / 512
    sc.params.wall_u_hi[2][0] = uw
    sc.params.detach_rate = 0.05 / 512
    sc.u0[0][:] = uw * (np.arange(sc.N) // (n * n) + 0.5) / n
    return sc


def cpu_leg_large(pfg, steps=200):
    """CPU leg of the large workload.  Checker: velocity and density L2 and
    platelet-centroid errors of the FP16-store device run against the FP64
    oracle on a 128^3 sub-box of the same physics (`steps` coupled steps),
    with an FP32-store device run of the same sub-box alongside: the Morse
    cutoff makes element trajectories sensitive to rounding, so the FP32
    figures show how much of the difference is not the storage format.
    Baseline: the oracle's rate on that same run."""
    import numpy as np
    po, threads = load_oracle()
    sc = subbox_scene(pfg)
    runs = {}
    for store in (pfg.STORE_F16, pfg.STORE_F32):
        sc.params.store = store
        with pfg.Solver(sc.params) as s:
            s.load(sc)
            s.step(steps)
            runs[store] = (s.fields(), s.markers())
    sc.params.store = pfg.STORE_F16
    o = po.Oracle(sc.params, scene=sc)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    orho, ox, oy, oz = po.fields(po.lattice(sc.params), o.f())
    OP = np.stack([o.marker(k) for k in ("x", "y", "z")], 1)
    o.close()
    body = sc.body
    npl = int(body.max()) + 1 if (body >= 0).any() else 0
    ocen = np.array([OP[body == b].mean(0) for b in range(npl)])
    X0 = np.stack([sc.x, sc.y, sc.z], 1).astype(np.float64)
    disp = np.array([np.linalg.norm(ocen[b] - X0[body == b].mean(0)) for b in range(npl)])
    den = np.sqrt((ox ** 2 + oy ** 2 + oz ** 2).sum())
    acc = {"case": f"128^3 sub-box of the large configuration, {npl} platelets, "
                   f"{sc.n_markers} markers, {steps} coupled steps, device vs FP64 oracle "
                   f"(FP32-store run of the same sub-box alongside)",
           "centroid_disp_mean": float(disp.mean()) if npl else None}
    for store, tag in ((pfg.STORE_F16, "f16"), (pfg.STORE_F32, "f32")):
        (rho, ux, uy, uz), mk = runs[store]
        P = np.stack([mk["x"], mk["y"], mk["z"]], 1).astype(np.float64)
        cen = np.array([P[body == b].mean(0) for b in range(npl)])
        dcen = np.linalg.norm(cen - ocen, axis=1)
        num = np.sqrt(((ux - ox) ** 2 + (uy - oy) ** 2 + (uz - oz) ** 2).sum())
        acc[tag] = {"rel_l2_u": float(num / den),
                    "rel_l2_rho": float(np.sqrt(((rho - orho) ** 2).sum() / (orho ** 2).sum())),
                    "max_abs_rho": float(np.abs(rho - orho).max()),
                    "centroid_err_max": float(dcen.max()) if npl else None,
                    "centroid_err_mean": float(dcen.mean()) if npl else None}
    cpu = {"value": round(sc.N * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": threads,
           "kind": "port",
           "sample": f"{steps} coupled steps of the 128^3 sub-box of the large workload, FP64 "
                     f"oracle, {threads} OpenMP threads ({dt:.1f} s)"}
    return cpu, acc


WORKLOADS = {
    "thrombus": ("SCENE_THROMBUS",
                 "thrombus in a tube: 512x256x256 lattice, solid-flag tube of radius 120 "
                 "along x, body force, 500 platelets x 1000 elements, fibrin 20k nodes / "
                 "30k links, FP32 storage"),
    "large": ("SCENE_LARGE",
              "large shear: 1024x512x512 lattice, Couette, 2000 platelets x 1000 "
              "elements, fibrin 80k nodes / 120k links, FP16 storage"),
}


def other_workload(pfg, name, K, W, cpu_leg=True):
    """Coupled steps and the LB kernel on a full-size configuration other than
    the headline one (same JSON schema; value = coupled MLUPS over all
    lattice cells)."""
    attr, desc = WORKLOADS[name]
    t0 = time.perf_counter()
    scene = pfg.Scene(getattr(pfg, attr))
    log(f"{name}: scene built in {time.perf_counter() - t0:.1f} s: {scene.N} cells, "
        f"{scene.n_markers} markers")
    store = scene.params.store
    s, dt = timed_run(pfg, scene, store, K, W, tune=name)
    info = s.step(0)
    dev_gb = s.device_bytes() / 1e9
    s.close()
    fluid = int((scene.flags == 0).sum()) if scene.flags is not None else scene.N
    bpc = BYTES_PER_CELL[store]
    lb_us = lb_kernel_us(pfg, scene, store, min(K, 200))
    # solid cells are skipped by the kernel: algorithmic bytes count fluid cells
    achieved = bpc * fluid / (lb_us * 1e-6) / 1e9
    mlups = scene.N * K / dt / 1e6
    dtype = {0: "f32", 1: "f16", 2: "bf16"}[store]
    cpu, acc = None, None
    if cpu_leg and name == "large":
        log("large: 128^3 sub-box against the FP64 oracle ...")
        cpu, acc = cpu_leg_large(pfg)
    return {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(dt * 1e3 / K, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded scene)",
        "config": {"workload": desc, "cells": scene.N, "fluid_cells": fluid,
                   "markers": scene.n_markers, "store": dtype},
        "coupled_steps_per_s": round(K / dt, 2),
        "fluid_mlups": round(fluid * K / dt / 1e6, 1),
        "device_gb": round(dev_gb, 2),
        "bonds_at_end": info["total_bonds"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": f"lb_fused_kernel<{dtype.upper()}>",
                     "bytes_per_cell": bpc, "lb_kernel_us": round(lb_us, 2)},
        "tuning": TUNING.get(name),
        "accuracy_vs_fp64": acc,
        "cpu_baseline": cpu,
    }


def batch_scenes(pfg, n=256, seed=1101):
    """The batched workload: n vessel segments, each axis U{32..96} cells, wall
    shear rate log-uniform in [10, 1000] /s mapped to a top-wall velocity of at
    most 0.05 (0.05 (rate/1000) (nz/96)), 1-8 platelets of 256 elements (fewer
    when they do not fit), 32 receptors per platelet, 512 wall ligands."""
    import numpy as np
    rng = np.random.default_rng(seed)
    scenes = []
    for k in range(n):
        nx, ny, nz = (int(v) for v in rng.integers(32, 97, size=3))
        rate = float(np.exp(rng.uniform(np.log(10.0), np.log(1000.0))))
        want = int(rng.integers(1, 9))
        sc = None
        for npl in range(want, 0, -1):
            try:
                sc = pfg.Scene(pfg.SCENE_BATCH, nx=nx, ny=ny, nz=nz, n_platelets=npl,
                               elements=256, receptors=32, ligands=512, seed=seed + k)
                break
            except pfg.PfgError:
                continue
        if sc is None:
            sc = pfg.Scene(pfg.SCENE_BATCH, nx=nx, ny=ny, nz=nz, n_platelets=-1,
                           receptors=-1, ligands=512, seed=seed + k)
        uw = 0.05 * (rate / 1000.0) * (nz / 96.0)
        sc.params.wall_u_hi[2][0] = uw
        sc.u0[0][:] = uw * (np.arange(sc.N) // (nx * ny) + 0.5) / nz
        scenes.append(sc)
    return scenes


def batch_workload(pfg, K, W, store):
    """256 lattices stepped as one batch, against the same 256 lattices as
    separate handles stepped one after the other."""
    t0 = time.perf_counter()
    scenes = batch_scenes(pfg)
    for sc in scenes:
        sc.params.store = store
    ncells = sum(sc.N for sc in scenes)
    nmark = sum(sc.n_markers for sc in scenes)
    log(f"batch: {len(scenes)} lattices, {ncells} cells, {nmark} markers "
        f"(built in {time.perf_counter() - t0:.1f} s)")
    b = pfg.Batch([sc.params for sc in scenes])
    b.load_scenes(scenes)
    if not NO_TUNE:
        TUNING["batch"] = b.autotune(16)
        log(f"autotune batch: {TUNING['batch']}")
    b.step(W)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(K)
    b.synchronize()
    dt_batch = time.perf_counter() - t0
    dev_gb = b.device_bytes() / 1e9
    b.close()
    solvers = []
    for sc in scenes:
        s = pfg.Solver(sc.params)
        s.load(sc)
        s.step(W)
        solvers.append(s)
    ks = max(1, K // 4)
    t0 = time.perf_counter()
    for s in solvers:
        s.step(ks)
    for s in solvers:
        s.synchronize()
    dt_sep = (time.perf_counter() - t0) * K / ks
    for s in solvers:
        s.close()
    dtype = {0: "f32", 1: "f16", 2: "bf16"}[store]
    return {
        "metric": METRIC, "value": round(ncells * K / dt_batch / 1e6, 1), "unit": "MLUPS",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": round(dt_batch * 1e3 / K, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded scenes)",
        "config": {"workload": "batch of 256 vessel segments, each axis U{32..96}, wall shear "
                               "log-U[10,1000]/s, 1-8 platelets x 256 elements, one handle",
                   "lattices": len(scenes), "cells": ncells, "markers": nmark, "store": dtype},
        "coupled_steps_per_s": round(K / dt_batch, 2),
        "separate_handles_ms_per_step": round(dt_sep * 1e3 / K, 4),
        "speedup_vs_separate": round(dt_sep / dt_batch, 2),
        "device_gb": round(dev_gb, 2),
        "tuning": TUNING.get("batch"),
        "roofline": None, "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU leg (FP64 oracle baseline and accuracy checks)")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--lb-only", choices=sorted(STORES), default=None)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--workload", choices=["shear", "batch"] + sorted(WORKLOADS),
                    default="shear", help="shear (default) is the headline configuration")
    ap.add_argument("--store", choices=sorted(STORES), default="f32",
                    help="storage of the batch workload")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-shape autotuner (keeps the automatic shapes)")
    args = ap.parse_args()
    global NO_TUNE
    NO_TUNE = args.no_tune
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    import platelet_fiber_gpu_amd as pfg
    if pfg.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)

    if args.copy_probe:
        import ctypes as C
        g = C.c_double()
        nbytes = 19 * 4 * 4194304
        rc = pfg.lib().pfg_copy_probe(0, nbytes, max(1, args.steps), C.byref(g))
        print(json.dumps({"copy_probe_gbs": round(g.value, 1), "bytes_each_way": nbytes,
                          "launches": args.steps, "status": rc}), flush=True)
        return

    K, W = args.steps, args.warmup
    if args.workload == "batch":
        print(json.dumps(batch_workload(pfg, K, W, STORES[args.store])), flush=True)
        return
    if args.workload != "shear":
        print(json.dumps(other_workload(pfg, args.workload, K, W, not args.no_cpu)),
              flush=True)
        return
    scene = pfg.Scene(pfg.SCENE_SHEAR)
    N = scene.N

    if args.lb_only:
        st = STORES[args.lb_only]
        s, dt = timed_run(pfg, scene, st, K, W, fluid_only=True)
        s.close()
        print(json.dumps({"lb_only": args.lb_only, "steps": K, "us_per_step": dt / K * 1e6,
                          "mlups": N * K / dt / 1e6}), flush=True)
        return

    log(f"workload 256x128x128 shear, {scene.n_markers} elements, "
        f"{scene.rec_tri.size} receptors, {scene.lig_x.size} ligands; K={K} W={W}")
    s32, dt32 = timed_run(pfg, scene, pfg.STORE_F32, K, W, tune="f32")
    info = s32.step(0)
    s32.close()
    ms_step = dt32 * 1e3 / K
    mlups = N * K / dt32 / 1e6
    log(f"fp32 coupled: {ms_step:.4f} ms/step, {mlups:.0f} MLUPS")

    launches = min(K, 500)
    lb32_us = lb_kernel_us(pfg, scene, pfg.STORE_F32, launches)
    achieved = BYTES_PER_CELL[0] * N / (lb32_us * 1e-6) / 1e9
    log(f"fp32 LB kernel: {lb32_us:.1f} us/launch, {achieved:.0f} GB/s")

    sl, dtl = timed_run(pfg, scene, pfg.STORE_F32, K, W, fluid_only=True)
    sl.close()
    lb32 = N * K / dtl / 1e6
    res16 = {}
    if not args.no_fp16:
        s16, dt16 = timed_run(pfg, scene, pfg.STORE_F16, K, W, tune="f16")
        s16.close()
        sl16, dtl16 = timed_run(pfg, scene, pfg.STORE_F16, K, W, fluid_only=True)
        sl16.close()
        lb16_us = lb_kernel_us(pfg, scene, pfg.STORE_F16, launches)
        res16 = {"coupled_mlups": round(N * K / dt16 / 1e6, 1),
                 "coupled_steps_per_s": round(K / dt16, 1),
                 "lb_only_mlups": round(N * K / dtl16 / 1e6, 1),
                 "lb_kernel_us": round(lb16_us, 2),
                 "lb_kernel_gbs": round(BYTES_PER_CELL[1] * N / (lb16_us * 1e-6) / 1e9, 1),
                 "speedup_coupled_vs_fp32": round(dt32 / dt16, 3),
                 "speedup_lb_only_vs_fp32": round(dtl / dtl16, 3),
                 "speedup_lb_kernel_vs_fp32": round(lb32_us / lb16_us, 3),
                 "rel_l2_u_vs_fp64": None,
                 "rel_l2_u_case": "64x32x32 channel, tau 0.8, F 1e-6, 300 steps; tol 1e-3"}
        log(f"fp16: {res16}")

    cpu = None
    if not args.no_cpu:
        log("cpu leg: FP64 oracle baseline and FP16 accuracy ...")
        cpu, err16, errb16 = cpu_leg_shear(pfg)
        if res16:
            res16["rel_l2_u_vs_fp64"] = err16
            res16["rel_l2_u_bf16_store_vs_fp64"] = errb16

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene)",
        "config": {"workload": "single platelet in shear, 256x128x128 D3Q19 Couette, "
                               "2048 elements, 512 receptors, 4096 ligands, FP32 storage",
                   "cells": N, "store": "f32"},
        "coupled_steps_per_s": round(K / dt32, 1),
        "lb_only_mlups_f32": round(lb32, 1),
        "bonds_at_end": info["total_bonds"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic(),
                     "kernel": "lb_fused_kernel<F32>", "bytes_per_cell": 152,
                     "lb_kernel_us": round(lb32_us, 2),
                     "mlups_kernel": round(N / (lb32_us * 1e-6) / 1e6, 1)},
        "fp16": res16,
        "tuning": TUNING,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
