#!/usr/bin/env python3
"""Benchmark of the coupled platelet-fibrin step on one MI355X.

Default workload: the 1024x512x512 configuration (268M lattice cells) --
D3Q19 Couette flow (wall velocity 0.05, tau 0.7), 2000 platelets of 1000 SEM
elements with 64 receptors each over 65536 wall ligands, a fibrin network of
80k nodes and 120k links, Bell bonds -- with the distributions stored as FP16
deviations (f_i - w_i) x 2^10 and collided in FP32.  A step is one full
coupled step: cell lists, membrane sums, bonds, forces, IB spreading, fused LB
collide-stream, IB interpolation and marker update.  Inputs are resident in
HBM before the timed region; the timed region is K hipGraph step launches
between two device synchronisations.  The same lattice then runs in FP32
storage for the FP16/FP32 ratios.

Prints ONE JSON line.  value = coupled-step MLUPS with FP16 storage (lattice
cells x steps / s).  roofline: the fused LB kernel (the dominant kernel),
algorithmic bytes 76 B/cell (19 x 2 B read + 19 x 2 B written) x 268435456
cells per launch over its mean duration, measured with HIP events on the
handle's stream around back-to-back launches of the same kernel.  traffic:
HBM bytes per launch from the committed rocprofv3 PMC summary of this
workload.  cpu_baseline and accuracy: the CPU leg on a 128^3 sub-box of the
same physics.

Other modes:
  --workload shear|thrombus|batch   the other configurations
  --lb-only f32|f16|bf16            K fused LB steps only (no markers) on the
                                    workload's lattice
  --copy-probe                      the library's float4 streaming copy
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "lb_fused_large1024_f16_traffic.txt")
PROFILE_STEPS = 40  # coupled steps whose LB launch is timed on its own
TRAFFIC_FILE_SHEAR = os.path.join(ROOT, "profiles", "lb_fused_shear256_traffic.txt")


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


def read_traffic(path, name="traffic_bytes_per_launch"):
    """HBM bytes per LB launch from a committed rocprofv3 PMC summary (the
    line `name = value`)."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        key, _, val = line.partition("=")
        if key.strip() == name:
            try:
                return float(val.split()[0])
            except (IndexError, ValueError):
                return None
    return None


def oracle_threads():
    """Host threads for the CPU leg: OMP_NUM_THREADS when it is set (the
    process's CPU share on a shared host), else every CPU of the process's
    affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, n)


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


def cpu_leg_large(pfg, horizons=(50, 200)):
    """CPU leg of the large workload, on a 128^3 sub-box of its physics
    (16 platelets, fibrin and wall ligands at its densities, hard Morse
    cutoff).  Checker: velocity and density L2 and platelet-centroid errors
    of FP16-store and FP32-store device runs against the FP64 oracle after
    each horizon, beside the distance between two FP64 runs whose initial
    positions differ by 1e-6 (how far the trajectory itself departs: the
    hard cutoff switches a pair's force a step apart in two runs that
    differ slightly).  Baseline: the oracle's rate on the same run."""
    import numpy as np
    po, threads = load_oracle()
    sc = pfg.Scene.large_subbox(128)
    npl = int(sc.plt_S0.size)
    body = sc.body
    dev = {}
    for store in (pfg.STORE_F16, pfg.STORE_F32):
        sc.params.store = store
        with pfg.Solver(sc.params) as s:
            s.load(sc)
            done = 0
            for h in horizons:
                s.step(h - done)
                done = h
                dev[(store, h)] = (s.fields(), s.markers())
    sc.params.store = pfg.STORE_F16
    L = po.lattice(sc.params)
    o = po.Oracle(sc.params, scene=sc)
    op = po.Oracle(sc.params, scene=sc)
    rng = np.random.default_rng(7)
    for k in ("x", "y", "z"):
        m = op.marker(k)
        m[:] = m + 1e-6 * rng.standard_normal(m.size)
    acc = {"case": f"128^3 sub-box of the large configuration (its hard Morse cutoff), "
                   f"{npl} platelets, {sc.n_markers} markers; device FP16 and FP32 "
                   f"storage against the FP64 oracle; 'floor': FP64 against FP64 with "
                   f"the initial positions perturbed by 1e-6", "tolerance_rel_l2_u": 1e-3}
    t_oracle, done = 0.0, 0
    for h in horizons:
        t0 = time.perf_counter()
        o.step(h - done)
        t_oracle += time.perf_counter() - t0
        op.step(h - done)
        done = h
        orho, ox, oy, oz = po.fields(L, o.f())
        den = np.sqrt((ox ** 2 + oy ** 2 + oz ** 2).sum())
        OP = np.stack([o.marker(k) for k in ("x", "y", "z")], 1)
        ocen = np.array([OP[body == b].mean(0) for b in range(npl)])
        entry = {}
        _, px, py, pz = po.fields(L, op.f())
        entry["floor_rel_l2_u"] = float(np.sqrt(((px - ox) ** 2 + (py - oy) ** 2 +
                                                 (pz - oz) ** 2).sum()) / den)
        for store, tag in ((pfg.STORE_F16, "f16"), (pfg.STORE_F32, "f32")):
            (rho, ux, uy, uz), mk = dev[(store, h)]
            P = np.stack([mk["x"], mk["y"], mk["z"]], 1).astype(np.float64)
            cen = np.array([P[body == b].mean(0) for b in range(npl)])
            num = np.sqrt(((ux - ox) ** 2 + (uy - oy) ** 2 + (uz - oz) ** 2).sum())
            entry[tag] = {"rel_l2_u": float(num / den),
                          "rel_l2_rho": float(np.sqrt(((rho - orho) ** 2).sum() /
                                                      (orho ** 2).sum())),
                          "centroid_err_max": float(np.linalg.norm(cen - ocen, axis=1).max())}
        acc[f"steps_{h}"] = entry
        log(f"sub-box {h} steps: {entry}")
    acc["oracle_near_cutoff_pair_visits"] = o.counters()["near_cutoff"]
    acc["oracle_ms_per_step"] = round(t_oracle * 1e3 / done, 1)
    o.close()
    op.close()
    # baseline: the same algorithm in FP32 on the host's cores (OpenMP),
    # about 10 s of coupled steps of the sub-box
    sc.params.store = pfg.STORE_F32
    port = po.CpuPort(sc.params, sc, threads)
    port.step(2)
    steps, t0 = 0, time.perf_counter()
    while steps < 8 or (time.perf_counter() - t0 < 10.0 and steps < 2000):
        port.step(4)
        steps += 4
    dt = time.perf_counter() - t0
    port.close()
    cpu = {"value": round(sc.N * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": threads,
           "kind": "port",
           "sample": f"{steps} coupled steps ({dt:.1f} s) of the large configuration's physics "
                     f"on a 128^3 sub-box at its densities ({sc.n_markers} markers), FP32 "
                     f"OpenMP port of the same algorithm on {threads} host threads; MLUPS is "
                     f"per lattice cell, as the headline value"}
    return cpu, acc


WORKLOADS = {
    "thrombus": ("SCENE_THROMBUS",
                 "thrombus in a tube: 512x256x256 lattice, solid-flag tube of radius 120 "
                 "along x, body force, 500 platelets x 1000 elements, fibrin 20k nodes / "
                 "30k links, FP32 storage"),
}

LARGE_DESC = ("1024x512x512 D3Q19 Couette (wall velocity 0.05, tau 0.7), 2000 platelets x "
              "1000 SEM elements (64 receptors each), 65536 wall ligands, fibrin 80k nodes / "
              "120k links, Bell bonds; FP16 deviation storage (x 2^10), FP32 collision")


def large_leg(pfg, scene, store, K, W):
    """One storage format on the 1024x512x512 lattice, one handle: autotuned
    coupled steps (K timed after W warm-up); the LB kernel's mean duration
    inside coupled steps (HIP events on the handle's stream around the LB
    launch of each of PROFILE_STEPS further steps, pfg_profile); then LB-only
    steps on the same lattice and the mean duration of back-to-back LB-only
    launches."""
    p = scene.params
    p.store = store
    t0 = time.perf_counter()
    s = pfg.Solver(p)
    s.load(scene)
    log(f"store {store}: handle loaded in {time.perf_counter() - t0:.1f} s, "
        f"{s.device_bytes() / 1e9:.1f} GB on the device")
    tag = {0: "f32", 1: "f16", 2: "bf16"}[store]
    if not NO_TUNE:
        TUNING[tag] = s.autotune(16)
        log(f"autotune {tag}: {TUNING[tag]}")
    s.step(W)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(K)
    s.synchronize()
    dt = time.perf_counter() - t0
    info = s.step(0)
    dev_gb = s.device_bytes() / 1e9
    n_prof = max(4, min(K, PROFILE_STEPS))
    _, lb_step_total = s.profile(n_prof)
    lb_step_ms = lb_step_total / n_prof
    s.set_fluid_only(True)
    kl = max(8, K // 2)
    s.step(8)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(kl)
    s.synchronize()
    dtl = time.perf_counter() - t0
    lb_ms = s.time_lb(max(10, min(K, 60)))
    s.close()
    N = scene.N
    bpc = BYTES_PER_CELL[store]
    gbs = bpc * N / (lb_ms * 1e-3) / 1e9
    gbs_step = bpc * N / (lb_step_ms * 1e-3) / 1e9
    out = {"coupled_ms_per_step": round(dt * 1e3 / K, 4),
           "coupled_mlups": round(N * K / dt / 1e6, 1),
           "coupled_steps_per_s": round(K / dt, 2),
           "lb_only_ms_per_step": round(dtl * 1e3 / kl, 4),
           "lb_only_mlups": round(N * kl / dtl / 1e6, 1),
           "lb_in_step_ms": round(lb_step_ms, 4),
           "lb_in_step_gbs": round(gbs_step, 1),
           "lb_in_step_roofline_frac": round(gbs_step / HBM_PEAK_GBS, 4),
           "lb_in_step_launches": n_prof,
           "lb_kernel_ms": round(lb_ms, 4),
           "lb_kernel_gbs": round(gbs, 1),
           "lb_kernel_roofline_frac": round(gbs / HBM_PEAK_GBS, 4),
           "lb_kernel_mlups": round(N / (lb_ms * 1e-3) / 1e6, 1),
           "bytes_per_cell": bpc,
           "bonds_at_end": info["total_bonds"],
           "bond_events": info["bond_events"],
           "device_gb": round(dev_gb, 2)}
    log(f"{tag}: {out}")
    return out


def large_workload(pfg, K, W, cpu_leg=True, fp32_leg=True):
    """The 1024x512x512 configuration, FP16 storage (the headline) and the
    same lattice in FP32 storage."""
    t0 = time.perf_counter()
    scene = pfg.Scene(pfg.SCENE_LARGE)
    log(f"large: scene built in {time.perf_counter() - t0:.1f} s: {scene.N} cells, "
        f"{scene.n_markers} markers; K={K} W={W}")
    r16 = large_leg(pfg, scene, pfg.STORE_F16, K, W)
    r32 = large_leg(pfg, scene, pfg.STORE_F32, K, W) if fp32_leg else None
    cpu, acc = None, None
    if cpu_leg:
        log("large: CPU leg (128^3 sub-box: CPU baseline and errors against FP64) ...")
        cpu, acc = cpu_leg_large(pfg)
    ratios = None
    if r32:
        ratios = {"coupled": round(r32["coupled_ms_per_step"] / r16["coupled_ms_per_step"], 3),
                  "lb_only": round(r32["lb_only_ms_per_step"] / r16["lb_only_ms_per_step"], 3),
                  "lb_kernel": round(r32["lb_kernel_ms"] / r16["lb_kernel_ms"], 3)}
    traffic = read_traffic(TRAFFIC_FILE, "traffic_bytes_per_launch")
    traffic_fo = read_traffic(TRAFFIC_FILE, "traffic_bytes_per_launch_fluid_only")
    return {
        "metric": METRIC, "value": r16["coupled_mlups"], "unit": "MLUPS", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": r16["coupled_ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene)",
        "config": {"workload": LARGE_DESC, "cells": scene.N, "markers": scene.n_markers,
                   "store": "f16", "arithmetic": "f32 (FP64 conservation and membrane sums)"},
        "coupled_steps_per_s": r16["coupled_steps_per_s"],
        "f16": r16,
        "f32": r32,
        "speedup_f16_over_f32": ratios,
        # the LB launch of the coupled step (it also reads and clears the IB
        # force and writes the velocity in the tiles this step's stencils
        # touch; those bytes are overhead, not algorithmic bytes)
        "roofline": {"bound": "hbm", "achieved": r16["lb_in_step_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": r16["lb_in_step_roofline_frac"],
                     "traffic": traffic, "kernel": "lb_fused_kernel<F16> in the coupled step",
                     "bytes_per_cell": BYTES_PER_CELL[1], "cells_per_launch": scene.N,
                     "lb_kernel_us": round(r16["lb_in_step_ms"] * 1e3, 1)},
        # the same kernel in LB-only steps, back to back (no IB tiles)
        "roofline_fluid_only": {
            "achieved": r16["lb_kernel_gbs"], "frac": r16["lb_kernel_roofline_frac"],
            "traffic": traffic_fo, "kernel": "lb_fused_kernel<F16>, LB-only launches",
            "lb_kernel_us": round(r16["lb_kernel_ms"] * 1e3, 1)},
        "roofline_f32": None if r32 is None else {
            "achieved": r32["lb_kernel_gbs"], "frac": r32["lb_kernel_roofline_frac"],
            "kernel": "lb_fused_kernel<F32>, LB-only launches",
            "bytes_per_cell": BYTES_PER_CELL[0], "target_frac": 0.70,
            "in_step_frac": r32["lb_in_step_roofline_frac"]},
        "tuning": TUNING,
        "accuracy_vs_fp64": acc,
        "cpu_baseline": cpu,
    }


def other_workload(pfg, name, K, W):
    """Coupled steps and the LB kernel on another full-size configuration
    (same JSON schema; value = coupled MLUPS over all lattice cells)."""
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
    fluid = scene.fluid_cells
    bpc = BYTES_PER_CELL[store]
    lb_us = lb_kernel_us(pfg, scene, store, min(K, 200))
    # solid cells are skipped by the kernel: algorithmic bytes count fluid cells
    achieved = bpc * fluid / (lb_us * 1e-6) / 1e9
    mlups = scene.N * K / dt / 1e6
    dtype = {0: "f32", 1: "f16", 2: "bf16"}[store]
    return {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(dt * 1e3 / K, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
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
        "cpu_baseline": None,
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
        sc.u0_z[0][:] = uw * (np.arange(nz) + 0.5) / nz
        scenes.append(sc)
    return scenes


def batch_workload(pfg, K, W, store, cpu_leg=True):
    """256 lattices stepped as one batch, against the same 256 lattices as
    separate handles, each created, stepped and destroyed in turn (one
    handle alive at a time)."""
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
    persistent = int((TUNING.get("batch") or {}).get("lb_persistent", 0))
    b.set_fluid_only(True)
    lb_ms_by = {}
    for form in (0, 1):  # one workgroup per x-block row; the persistent queue launch
        b.set_lb_persistent(form)
        b.step(8)
        lb_ms_by[form] = b.time_lb(max(10, min(K, 60)))
    b.close()
    lb_ms = lb_ms_by[persistent]
    bpc = BYTES_PER_CELL[store]
    gbs = bpc * ncells / (lb_ms * 1e-3) / 1e9
    ks = max(1, K // 4)
    dt_sep = 0.0
    for sc in scenes:
        with pfg.Solver(sc.params) as s:
            s.load(sc)
            s.step(W)
            s.synchronize()
            t1 = time.perf_counter()
            s.step(ks)
            s.synchronize()
            dt_sep += time.perf_counter() - t1
    dt_sep *= K / ks
    cpu = None
    if cpu_leg:
        po, threads = load_oracle()
        sample = scenes[:16]
        cells, t_cpu = 0, 0.0
        for sc in sample:
            sc.params.store = pfg.STORE_F32
            port = po.CpuPort(sc.params, sc, threads)
            port.step(1)
            t1 = time.perf_counter()
            port.step(20)
            t_cpu += time.perf_counter() - t1
            cells += sc.N * 20
            port.close()
            sc.params.store = store
        cpu = {"value": round(cells / t_cpu / 1e6, 3), "unit": "MLUPS", "cores": threads,
               "kind": "port",
               "sample": f"20 coupled steps of each of the first 16 lattices of the batch "
                         f"({t_cpu:.1f} s), FP32 OpenMP port of the same algorithm on "
                         f"{threads} host threads, one lattice after the other"}
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
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": ("lb_batch_queue_kernel" if persistent else "lb_batch_kernel")
                     + f"<{dtype.upper()}>", "bytes_per_cell": bpc,
                     "cells_per_launch": ncells, "lb_kernel_us": round(lb_ms * 1e3, 1),
                     "lb_kernel_us_per_row_launch": round(lb_ms_by[0] * 1e3, 1),
                     "lb_kernel_us_persistent": round(lb_ms_by[1] * 1e3, 1)},
        "cpu_baseline": cpu,
    }


def shear_workload(pfg, K, W, cpu_leg=True, fp16_leg=True):
    """The 256x128x128 single platelet in shear (one 2048-element platelet,
    512 receptors, 4096 wall ligands), FP32 storage, with the FP16 figures of
    the same workload alongside."""
    scene = pfg.Scene(pfg.SCENE_SHEAR)
    N = scene.N
    log(f"workload 256x128x128 shear, {scene.n_markers} elements, "
        f"{scene.rec_tri.size} receptors, {scene.lig_x.size} ligands; K={K} W={W}")
    s32, dt32 = timed_run(pfg, scene, pfg.STORE_F32, K, W, tune="f32")
    info = s32.step(0)
    s32.close()
    ms_step = dt32 * 1e3 / K
    mlups = N * K / dt32 / 1e6
    launches = min(K, 500)
    lb32_us = lb_kernel_us(pfg, scene, pfg.STORE_F32, launches)
    achieved = BYTES_PER_CELL[0] * N / (lb32_us * 1e-6) / 1e9
    sl, dtl = timed_run(pfg, scene, pfg.STORE_F32, K, W, fluid_only=True)
    sl.close()
    res16 = {}
    if fp16_leg:
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
                 "speedup_lb_kernel_vs_fp32": round(lb32_us / lb16_us, 3)}
    cpu = None
    if cpu_leg:
        log("cpu leg: FP64 oracle baseline and FP16 accuracy ...")
        cpu, err16, errb16 = cpu_leg_shear(pfg)
        if res16:
            res16["rel_l2_u_vs_fp64"] = err16
            res16["rel_l2_u_bf16_store_vs_fp64"] = errb16
            res16["rel_l2_u_case"] = "64x32x32 channel, tau 0.8, F 1e-6, 300 steps; tol 1e-3"
    return {
        "metric": METRIC, "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(ms_step, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene)",
        "config": {"workload": "single platelet in shear, 256x128x128 D3Q19 Couette, "
                               "2048 elements, 512 receptors, 4096 ligands, FP32 storage",
                   "cells": N, "store": "f32"},
        "coupled_steps_per_s": round(K / dt32, 1),
        "lb_only_mlups_f32": round(N * K / dtl / 1e6, 1),
        "bonds_at_end": info["total_bonds"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": read_traffic(TRAFFIC_FILE_SHEAR),
                     "kernel": "lb_fused_kernel<F32>", "bytes_per_cell": 152,
                     "lb_kernel_us": round(lb32_us, 2)},
        "fp16": res16,
        "tuning": TUNING,
        "cpu_baseline": cpu,
    }


def lb_only(pfg, workload, store, K, W):
    """K fused LB steps on the workload's lattice without markers."""
    if workload == "shear":
        scene = pfg.Scene(pfg.SCENE_SHEAR, n_platelets=-1, receptors=-1, ligands=-1)
    elif workload == "large":
        scene = pfg.Scene(pfg.SCENE_LARGE, n_platelets=-1, receptors=-1, ligands=-1,
                          fibrin_nodes=-1)
    else:
        raise SystemExit(f"--lb-only runs on the shear or large lattice, not {workload}")
    s, dt = timed_run(pfg, scene, store, K, W, fluid_only=True)
    ms = s.time_lb(max(1, min(K, 60)))
    s.close()
    return {"lb_only": store, "workload": workload, "cells": scene.N, "steps": K,
            "ms_per_step": round(dt / K * 1e3, 4), "mlups": round(scene.N * K / dt / 1e6, 1),
            "lb_kernel_ms": round(ms, 4),
            "lb_kernel_gbs": round(BYTES_PER_CELL[store] * scene.N / (ms * 1e-3) / 1e9, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU leg (CPU baseline and accuracy against FP64)")
    ap.add_argument("--no-f32", action="store_true",
                    help="large workload: skip the FP32-store leg")
    ap.add_argument("--no-fp16", action="store_true", help="shear workload: FP32 only")
    ap.add_argument("--lb-only", choices=sorted(STORES), default=None)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--workload", choices=["large", "shear", "batch"] + sorted(WORKLOADS),
                    default="large", help="large (default) is the headline configuration")
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
    if args.lb_only:
        print(json.dumps(lb_only(pfg, args.workload, STORES[args.lb_only], K, W)), flush=True)
    elif args.workload == "large":
        print(json.dumps(large_workload(pfg, K, W, not args.no_cpu, not args.no_f32)),
              flush=True)
    elif args.workload == "shear":
        print(json.dumps(shear_workload(pfg, K, W, not args.no_cpu, not args.no_fp16)),
              flush=True)
    elif args.workload == "batch":
        print(json.dumps(batch_workload(pfg, K, W, STORES[args.store], not args.no_cpu)),
              flush=True)
    else:
        print(json.dumps(other_workload(pfg, args.workload, K, W)), flush=True)


if __name__ == "__main__":
    main()
