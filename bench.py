#!/usr/bin/env python3
"""Throughput of the MLS-MPM substep on one MI355X.

Workload (default): the liquid dam break -- 1,000,000 particles uniform in
the box [0.05,0.45]x[0.05,0.65]x[0.05,0.95], grid N=128, dt=5e-5, weakly
compressible liquid K=4e4, rho=1000, gravity -9.8 (seeded synthetic input,
mpm_amd.scenes.dam_break). A step is one substep of the engine (grid update +
fused G2P/P2G, with the periodic cell sort) over all particles, with the
particle state resident in HBM. The default run times 2,000 substeps, the
dam break's length. Prints one JSON line:

  value        particle-substeps/s (P / seconds per substep), whole job
  roofline     the dominant kernel (largest share of a substep's time, the
               sort amortised over its interval): its algorithmic bytes per
               launch divided by its mean launch time, against 8 TB/s HBM.
               Launch times are HIP events the engine records on its launch
               stream around one substep in every sort interval of the timed
               region; when that gives fewer than 8 launches (a short run),
               the substeps right after the region are timed one by one to
               make up 8. `traffic` is the rocprofv3-counted HBM bytes per
               launch of that kernel (profiles/traffic_<workload>.json)
  cpu_baseline the FP32 CPU oracle (a C++/OpenMP restatement of the
               reference step) on a bounded sample of the same workload, on
               the OpenMP thread budget and on one thread, plus the 10,000-
               particle elastic sphere; `accuracy` holds the engine's error
               against the FP64 oracle over the workload's first 10 substeps
  snow         the snow scene (500,000 particles in two stacked spheres, N=64,
               dt=1e-4, snow plasticity and hardening): its 1,000 substeps
               after tuning, particle-substeps/s and fraction of the HBM
               roofline (its error against the FP64 oracle sits in
               cpu_baseline's accuracy)
  batched      64 independent scenes advanced one after the other vs. as one
               batch (HIP graph replay): seconds and speedup
  adjoint      the differentiable elastic cube (32,768 particles, N=32, 200
               substeps, checkpoint every 20): forward and backward wall time
               and their ratio

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
TIMING_EVERY = 16
MIN_TIMED_LAUNCHES = 8


def alg_bytes(phase: str, P: int, N: int) -> int:
    """Algorithmic bytes of one launch of a kernel phase (dense FP32 grid):
    g2p2g   G2P reads x,F,Jp,mat (53 B) and writes x,v,C,F,Jp (100 B); the
            fused P2G uses the state it just computed, whose algorithmic read
            (101 B) is counted as in the three-phase step: 254 B/particle;
            grid: gather of the velocities and write of the P2G result,
            2 x 16 B/node
    grid_update  (the tile reduction with the grid update) read of the P2G
            result and write of the updated node: 2 x 16 B/node
    """
    if phase == "g2p2g":
        return 254 * P + 32 * N ** 3
    if phase == "grid_update":
        return 32 * N ** 3
    return 0


def load_traffic(workload: str, kernel: str):
    """HBM bytes per launch of the kernel from the committed rocprofv3 PMC
    summary (profiles/traffic_<workload>.json: FETCH_SIZE + WRITE_SIZE of the
    kernel, each scaled by the copy-kernel calibration), or None."""
    f = ROOT / "profiles" / f"traffic_{workload}.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        k = d["kernels"].get(kernel)
        return None if k is None else int(k["hbm_bytes_per_launch"])
    except Exception:
        return None


def host_cpu() -> dict:
    """The host CPU as the OS reports it: model name, logical CPUs, the CPUs
    this process may run on, and the OpenMP thread budget (OMP_NUM_THREADS)."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return {"model": model, "logical_cpus": os.cpu_count(), "cpus_allowed": affinity,
            "omp_num_threads": int(omp) if omp.isdigit() else None}


def cpu_rate(scene, steps: int, threads: int, repeats: int, warm: int = 1) -> dict:
    """Particle-substeps/s of the FP32 CPU oracle on `scene`: median of
    `repeats` timed runs of `steps` substeps each, after `warm` warm-up
    substeps; the state carries over from run to run."""
    import oracle
    st = scene.state()
    oracle.time_advance_f32(scene.params, scene.materials, st, warm, threads)  # warm-up
    secs = sorted(oracle.time_advance_f32(scene.params, scene.materials, st, steps, threads)
                  for _ in range(repeats))
    med = secs[len(secs) // 2]
    return {"value": scene.n * steps / med, "threads": threads, "substeps": steps,
            "repeats": repeats, "median_s": med, "workload": scene.name, "particles": scene.n,
            "grid_N": scene.params.N}


def fp64_check(scene, options: dict, steps: int, threads: int) -> dict:
    """Error of the engine's first `steps` substeps of `scene` (run with the
    bench's options) against the FP64 CPU oracle of the same step, with the
    FP32 oracle's own error beside it. Per field: the norm-relative error
    max|a - b| / max|b| and the element-wise one max |a - b| / max(|b|,
    1e-3 max|b|) (C against the scale max(|C|, |v| N))."""
    import numpy as np
    import oracle
    from mpm_amd import Engine

    st = scene.state()
    eng = Engine(scene.params)
    eng.set_materials(scene.materials)
    for k, v in options.items():
        eng.set_option(k, v)
    eng.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    eng.upload()
    eng.advance(steps)
    got = eng.download()
    eng.close()
    ref64 = oracle.advance(scene.params, scene.materials, st, steps, np.float64, threads=threads)
    ref32 = oracle.advance(scene.params, scene.materials, st, steps, np.float32, threads=threads)
    N = scene.params.N

    def errs(a, b):
        out = {}
        vmax = float(np.abs(b["v"]).max())
        for f in ("x", "v", "C", "F", "Jp"):
            A = np.asarray(a[f], np.float64)
            B = np.asarray(b[f], np.float64)
            scale = max(float(np.abs(B).max()), vmax * N if f == "C" else 0.0)
            diff = np.abs(A - B)
            out[f] = {"norm_rel": float(diff.max() / scale) if scale > 0 else float(diff.max()),
                      "elem_rel": float((diff / np.maximum(np.abs(B), 1e-3 * scale)).max())
                      if scale > 0 else float(diff.max())}
        return out

    return {"substeps": steps, "workload": scene.name, "gpu_vs_fp64": errs(got, ref64),
            "fp32_oracle_vs_fp64": errs(ref32, ref64)}


def cpu_rate_lines(workload: str, cpu_steps: int, threads: int, c1_steps: int = 100) -> list:
    """The CPU baseline's timing lines: the workload on `threads` threads (5
    runs of a fifth of `cpu_steps`) and on one; the 10,000-particle elastic
    sphere (N=16) on both, median of 5 runs of `c1_steps` substeps after 10
    warm-up substeps."""
    from mpm_amd import scenes
    scene = getattr(scenes, workload)()
    c1 = scenes.elastic_sphere()
    return [cpu_rate(scene, max(1, cpu_steps // 5), threads, 5), cpu_rate(scene, 1, 1, 3),
            cpu_rate(c1, c1_steps, threads, 5, warm=10), cpu_rate(c1, c1_steps, 1, 5, warm=10)]


def cpu_rate_lines_bound(workload: str, cpu_steps: int, threads: int, c1_steps: int = 100):
    """cpu_rate_lines in a child process whose OpenMP threads are bound to
    the cores they start on, one per core (OMP_PROC_BIND=close and
    OMP_PLACES=cores, unless the caller set them). The runtime binds when it
    loads, the calling thread included, so the binding stays out of this
    process and its HIP threads."""
    import subprocess
    env = dict(os.environ)
    env.setdefault("OMP_PROC_BIND", "close")
    env.setdefault("OMP_PLACES", "cores")
    r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-rates", workload,
                        "--cpu-steps", str(cpu_steps), "--cpu-threads", str(threads),
                        "--cpu-c1-steps", str(c1_steps)],
                       env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline child failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1]), f"{env['OMP_PROC_BIND']}, places {env['OMP_PLACES']}"


def cpu_baseline_leg(scene, options: dict, cpu_steps: int):
    """The reported CPU baseline (the FP32 C++/OpenMP oracle, `omp atomic`
    scatter) on the host of the GPU: the bench workload on every thread of the
    OpenMP budget, and on one thread; the 10,000-particle elastic sphere (N=16,
    100 substeps) on both; and, as the check that the numbers time the same
    step, the engine's first 10 substeps of the workload against the FP64
    oracle."""
    from mpm_amd import scenes

    cpu = host_cpu()
    threads = cpu["omp_num_threads"] or cpu["cpus_allowed"] or 1
    lines, cpu["omp_proc_bind"] = cpu_rate_lines_bound(scene.name, cpu_steps, threads)
    main = lines[0]
    snow = scenes.snow_spheres()
    sample = (f"{scene.name}: all {scene.n} particles, N={scene.params.N}, median of 5 runs of "
              f"{main['substeps']} FP32 substeps after 1 warm-up, C++/OpenMP oracle (omp atomic "
              f"scatter), {threads} threads ({main['median_s']:.2f} s per run); CPU "
              f"{cpu['model']}, {cpu['cpus_allowed']} CPUs allowed of {cpu['logical_cpus']}")
    return {
        "value": main["value"], "unit": "particle-substeps/s", "cores": threads, "kind": "port",
        "sample": sample, "host": cpu,
        "lines": lines,
        "accuracy": fp64_check(scene, options, 10, threads),
        "accuracy_snow": fp64_check(snow, {}, 10, threads),
    }


def snow_leg(steps: int, warm: int = 64):
    """The snow scene (mpm_amd.scenes.snow_spheres: 500,000 particles, N=64):
    `steps` substeps after the library's tuner and `warm` warm-up substeps,
    wall time between device synchronisations."""
    import numpy as np
    import torch
    from mpm_amd import Engine, scenes

    sc = scenes.snow_spheres()
    st = sc.state()
    eng = Engine(sc.params)
    eng.set_materials(sc.materials)
    eng.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    eng.upload()
    eng.autotune(trial_substeps=64)
    eng.advance(warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.advance(steps)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    out = eng.download()
    opts = {k: eng.get_option(k) for k in ("sort_every", "chunk", "radix_bits", "combine")}
    eng.close()
    b_alg = scenes.algorithmic_bytes(sc.n, sc.params.N)
    return {
        "workload": f"{sc.name}: {sc.n} particles, N={sc.params.N}, dt={sc.params.dt}",
        "substeps": steps, "warmup": warm, "seconds": t, "us_per_substep": 1e6 * t / steps,
        "particle_substeps_per_s": sc.n * steps / t,
        "algorithmic_bytes_per_substep": b_alg,
        "fraction_of_hbm_roofline": b_alg * steps / t / (HBM_PEAK_GBS * 1e9),
        "options": opts,
        "finite": bool(all(np.isfinite(out[f]).all() for f in ("x", "v", "F"))),
    }


def batched_leg(warm: int, steps: int):
    """The batched workload (64 independent scenes, mpm_amd.scenes.batched_scenes):
    seconds for `steps` substeps of every scene advanced one scene after the
    other (each handle's own mpm_advance, with the settings its own tuner
    picks), against the same substeps through one batch (HIP graph replay of
    multi-scene launches: one per phase for all scenes) under the settings
    the batch tuner picks. Both are warmed by `warm` substeps after tuning."""
    import torch
    from mpm_amd import Batch, Engine, scenes

    def make(sc):
        st = sc.state()
        e = Engine(sc.params)
        e.set_materials(sc.materials)
        e.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
        e.upload()
        return e

    scs = scenes.batched_scenes()
    seq = [make(sc) for sc in scs]
    for e in seq:
        e.autotune(trial_substeps=64)
        e.advance(warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e in seq:
        e.advance(steps)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    for e in seq:
        e.close()
    bat = [make(sc) for sc in scs]
    b = Batch(bat)
    tuned = b.autotune(trial_substeps=64)
    b.advance(warm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.advance(steps)
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t0
    launches = b.get_option("graph_launches")
    fused = b.get_option("fused_captures")
    b.close()
    for e in bat:
        e.close()
    P = sum(sc.n for sc in scs)
    # algorithmic bytes of one substep of every scene (254 B per particle, 64 B
    # per node of its own grid)
    b_alg = sum(scenes.algorithmic_bytes(sc.n, sc.params.N) for sc in scs)
    return {
        "workload": f"{len(scs)} independent scenes, {P} particles, N in {{32,48,64}}",
        "substeps": steps, "warmup": warm,
        "sequential_s": t_seq, "batched_s": t_bat, "speedup": t_seq / t_bat,
        "batched_particle_substeps_per_s": P * steps / t_bat,
        "algorithmic_bytes_per_substep": b_alg,
        "batched_fraction_of_hbm_roofline": b_alg * steps / t_bat / (HBM_PEAK_GBS * 1e9),
        "sequential_fraction_of_hbm_roofline": b_alg * steps / t_seq / (HBM_PEAK_GBS * 1e9),
        "batch_tuner_choice": tuned["choice"],
        "batch_tuner_trials": tuned["trials"],
        "graph_launches": launches,
        "multi_scene_cycles_recorded": fused,
    }


def adjoint_leg(repeats: int = 3):
    """The differentiable elastic cube (mpm_amd.scenes.diff_cube: 32,768
    particles, N=32, 200 substeps, checkpoint every 20): wall time of the
    forward rollout and of the backward (segment recompute + adjoint kernels),
    best of `repeats`."""
    import torch
    from mpm_amd import scenes
    from mpm_amd.diff import DiffStepper, lame_from_logE, pack_state, rollout

    c = scenes.diff_cube()
    n = c.x.shape[0]
    dev = torch.device("cuda")
    stp = DiffStepper(c.params, c.particle_mass, c.particle_volume, n)
    x = torch.from_numpy(c.x).to(dev)
    target = torch.tensor(c.target, dtype=torch.float32, device=dev)
    best = None
    for r in range(repeats + 1):  # the first round warms up
        v0 = torch.zeros(3, device=dev, requires_grad=True)
        logE = torch.from_numpy(c.logE).to(dev).requires_grad_(True)
        mu, lam = lame_from_logE(logE, c.nu)
        st0 = pack_state(x, v0.expand(n, 3))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = rollout(stp, st0, mu, lam, c.substeps, 20)
        loss = ((st[0:3].mean(1) - target) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if r > 0 and (best is None or (t2 - t1) / (t1 - t0) < best[1] / best[0]):
            best = (t1 - t0, t2 - t1)
    fwd, bwd = best
    eng_us = engine_cube_forward_us(c)
    return {
        "workload": f"differentiable elastic cube, {n} particles, N={c.params.N}, "
                    f"{c.substeps} substeps, checkpoint every 20",
        "forward_ms": fwd * 1e3, "backward_ms": bwd * 1e3,
        "forward_us_per_substep": fwd * 1e6 / c.substeps,
        "backward_us_per_substep": bwd * 1e6 / c.substeps,
        "bwd_fwd_ratio": bwd / fwd,
        "engine_forward_us_per_substep": eng_us,
        "bwd_over_engine_forward": bwd * 1e6 / c.substeps / eng_us,
    }


def engine_cube_forward_us(c, steps: int = 400) -> float:
    """Microseconds per substep of the same cube through the engine's plain
    tiled step (mpm_advance; the fixed-corotated model with J = det F, compat
    FIXED, one E = 1e4 for every particle, since the engine's materials are
    per kind, not per particle), after 48 warm-up substeps."""
    import numpy as np
    import torch
    from mpm_amd import MATERIAL_FIXED_COROTATED, Engine, material_as_numpy

    mat = material_as_numpy(MATERIAL_FIXED_COROTATED, c.particle_volume, 1000.0, E=1e4, nu=c.nu)
    mat.particle_mass = c.particle_mass
    e = Engine(c.params)
    e.set_materials([mat])
    e.add_particles(c.x, np.zeros(c.x.shape[0], np.uint8))
    e.upload()
    e.advance(48)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e.advance(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    e.close()
    return (t1 - t0) * 1e6 / steps


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="dam_break", choices=["dam_break", "snow_spheres", "elastic_sphere"])
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="substeps of the CPU baseline's main line (5 timed runs of a fifth each)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-snow", action="store_true")
    ap.add_argument("--snow-steps", type=int, default=1000)
    ap.add_argument("--batched-steps", type=int, default=500)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    # (the CPU baseline's child process: its timing lines as one JSON line)
    ap.add_argument("--cpu-rates", default="", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-threads", type=int, default=1, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-c1-steps", type=int, default=100, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_rates:
        print(json.dumps(cpu_rate_lines(args.cpu_rates, args.cpu_steps, args.cpu_threads, args.cpu_c1_steps)),
              flush=True)
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: the engine advances one simulation per GPU and this "
              "build has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        return 2

    import torch
    import mpm_amd
    from mpm_amd import Engine, scenes

    scene = {"dam_break": scenes.dam_break, "snow_spheres": scenes.snow_spheres,
             "elastic_sphere": scenes.elastic_sphere}[args.workload]()
    P, N = scene.n, scene.params.N
    st = scene.state()
    eng = Engine(scene.params)
    eng.set_materials(scene.materials)
    eng.add_particles(st["x"], st["material"], st["v"], st["C"], st["F"], st["Jp"])
    eng.upload()
    eng.advance(args.warmup)
    tuned = None
    if not args.no_autotune:
        # the library's autotuner picks the sort interval and work-item size
        # for this scene (from a snapshot; the state is put back)
        tuned = eng.autotune(trial_substeps=64)
        # back to the steady state under the chosen options (the tuner
        # restored the state as it was, with the items to rebuild)
        eng.advance(max(args.warmup, 2 * eng.get_option("sort_every")))
    torch.cuda.synchronize()

    # timed region: events on the launch stream bracket the kernel phases of
    # one substep in every k (k = the sort interval), rotating through the k
    # positions of a k-substep cycle; timing more substeps would slow the
    # wall clock (each event pair waits for its kernel to drain)
    sort_every = eng.get_option("sort_every") or TIMING_EVERY
    eng.set_option("timing", sort_every)
    eng.timing_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.advance(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    in_region = {ph: eng.timing(ph)[1] for ph in ("g2p2g", "grid_update")}
    # a short run times few launches: the substeps right after the region,
    # every one timed, bring each per-substep phase to MIN_TIMED_LAUNCHES
    top_up = max(0, MIN_TIMED_LAUNCHES - min(in_region.values()))
    if top_up:
        eng.set_option("timing", 1)
        eng.advance(top_up)
        torch.cuda.synchronize()
    eng.set_option("timing", 0)
    phases = {}
    for ph in ("g2p2g", "grid_update", "sort", "p2g", "g2p", "clear"):
        ms, n = eng.timing(ph)
        if n:
            phases[ph] = {"mean_us": 1e3 * ms / n, "launches": n}
    kernel_timing = {"timed_in_region": in_region, "timed_after_region": top_up}
    eng_opts = {k: eng.get_option(k) for k in ("sort_every", "chunk", "radix_bits")}
    out = eng.download()
    eng.close()
    finite = bool(torch.isfinite(torch.from_numpy(out["x"])).all())

    sec_per_step = (t1 - t0) / args.steps
    # the dominant kernel: the largest share of a substep's time, the sort
    # counted once per sort interval, every other phase once per substep
    per_substep = {ph: d["mean_us"] / (sort_every if ph == "sort" else 1) for ph, d in phases.items()}
    dom = max(per_substep, key=per_substep.get)
    mean_s = phases[dom]["mean_us"] * 1e-6
    achieved = alg_bytes(dom, P, N) / mean_s / 1e9
    kernel_name = {"g2p2g": "k_g2p2g_tiled", "grid_update": "k_tile_reduce"}.get(dom, dom)

    result = {
        "metric": "particle-substeps/s at 1M particles N=128; batched speedup; bwd/fwd ratio",
        "value": P / sec_per_step,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scene, no checkpoint)",
        "config": {
            "workload": scene.name, "particles": P, "grid_N": N, "dt": scene.params.dt,
            "material": {"dam_break": "weakly compressible liquid K=4e4 rho=1000",
                         "snow_spheres": "snow E=1.4e5 nu=0.2 rho=400 xi=10",
                         "elastic_sphere": "elastic (snow model, clamps [0,1e30]) E=1e4 nu=0.3"}[args.workload],
            "substep_kernels": phases,
            "kernel_timing": kernel_timing,
            "options": {"sort_every": eng_opts["sort_every"], "chunk": eng_opts["chunk"],
                        "radix_bits": eng_opts["radix_bits"], "autotuned": tuned is not None},
            "algorithmic_bytes_per_substep": scenes.algorithmic_bytes(P, N),
            "substep_fraction_of_hbm_roofline":
                scenes.algorithmic_bytes(P, N) / sec_per_step / (HBM_PEAK_GBS * 1e9),
            "finite": finite,
            "device": mpm_amd._lib.device_info(),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_name,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(scene.name, kernel_name),
        },
        "cpu_baseline": None,
        "snow": None,
        "batched": None,
        "adjoint": None,
    }
    if not args.no_snow:
        result["snow"] = snow_leg(args.snow_steps)
        print(f"snow: {result['snow']}", file=sys.stderr, flush=True)
    if not args.no_batched:
        result["batched"] = batched_leg(48, args.batched_steps)
        print(f"batched: {result['batched']}", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        result["adjoint"] = adjoint_leg()
        print(f"adjoint: {result['adjoint']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline_leg(scene, eng_opts, args.cpu_steps)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
