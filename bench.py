"""Benchmark of the coupled LBM step on one MI355X.

Workload: the 256^3 single scene of config 1 (liquid 40%, perturbed interface,
falling sphere r=24), float32 populations, FP64 arithmetic; one coupled step
= sim.step; absorbDeadCells; sphere step.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--size 256]

MLUPS counts all nx*ny*nz cells, walls included, over the timed steps of a
continuing run (lbm_step_live: no per-call prologue or closing stream-only
step inside the timed region, every step complete; `reference_state_call`
gives the same steps as one call from and to the reference's state).  Launch
shapes come from the autotuner; no timing events in the timed region.  The
roofline object is for the dominant kernel (stream + macro + next collide),
timed with HIP events on the library's stream over a further call.  For
float32 at 256^3 that is the row march with the band box's launch beside it
(k1_march computes phi(t+1) itself: 280 B per cell, the 27+7 populations read
and written, phi(t) read and phi(t+1) written); otherwise the bulk and
band-row launches of the stream kernel (281 B per cell in float32: the
populations, phi(t), phi(t+1) and the tag read; 145 B with 16-bit
populations), after a phase pass of their own.  Bulk cells store rho and u
only at the last step of a call.  The whole-step fraction uses 297 B (161 B)
per cell (populations, phi read and written, rho/u written, the tag).
`traffic` is the measured HBM bytes per launch from the committed counter
profiles (rocprofv3 FETCH_SIZE/WRITE_SIZE calibrated on a copy kernel), for
the 256^3 scene only.

The default line also carries the FP16-deviation form on the same scene
(`f16`, with its ratio to float32), config 2 stepped as one batch in both
forms (`batch`, with the ratio) and the serving loop (config 4: Poisson
arrivals into the request loop) in both forms for --serve-seconds each
(`serve`: served MLUPS, frame latency and time-to-first-frame percentiles).
Other workloads: --workload batch (config 2) and --workload serve.
The CPU baseline is the oracle (a single-threaded C restatement of the
reference) on the same scene for a few steps."""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "claude-vibe-code-raytracing_amd")
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402

import lbm  # noqa: E402
import scenegen  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per lattice update of the dominant kernel and of the whole
# step, per population storage: float32 (4 B per population) or a 16-bit
# deviation (2 B).  Kernel: 34 populations read and written + tag + phi(t) +
# phi(t+1); step: 34 populations read and written + phi read and written +
# rho/u written + tag.
KERNEL_BYTES_PER_LU = {"f32": 281, "f16": 145, "bf16": 145}
# The float32 row march (lbm_set_stream_split 2) computes phi(t+1) itself: it
# reads the 27 + 7 populations and phi(t), writes the 27 + 7 populations and
# phi(t+1) (the tag only inside the band box).
MARCH_BYTES_PER_LU = 280
MARCH_KERNEL = "k1_march + k1_stream_collide_box"
STEP_BYTES_PER_LU = {"f32": 297, "f16": 161, "bf16": 161}
# Populations alone (27 + 7 planes read and written): the pattern probe's bytes.
PATTERN_BYTES_PER_LU = {"f32": 272, "f16": 136, "bf16": 136}
TRAFFIC_JSON = {"f32": os.path.join(ROOT, "profiles", "stream_collide_256cube_traffic.json"),
                "f16": os.path.join(ROOT, "profiles", "stream_collide_256cube_f16_traffic.json")}


def measured_traffic(n, storage):
    """HBM bytes per launch of the dominant kernel from the committed counter
    profile, when it was taken on the same grid and storage."""
    try:
        with open(TRAFFIC_JSON[storage]) as fh:
            t = json.load(fh)
    except (KeyError, OSError, ValueError):
        return None
    return t["traffic_bytes"] if t.get("cells") == n ** 3 else None
STORAGE_NAME = {"f32": "float32", "f16": "fp16-deviation", "bf16": "bf16-deviation"}
# The dominant kernel per storage: float32 runs one cell per lane, the 16-bit
# codecs two (aligned 4-byte pairs).
STREAM_KERNEL = {"f32": "k1_stream_collide", "f16": "k1_stream_collide2",
                 "bf16": "k1_stream_collide2"}
ARITH = {"f32": "f64", "f16": "f32", "bf16": "f32"}
METRIC = "coupled MLUPS + %HBM peak at 256^3; served scenes MLUPS, p99 frame latency"


def build_scene(n, seed):
    params, yw, sph = scenegen.config1(n, n, n, seed)
    sim = lbm.Sim(n, n, n)
    sim.set_params(**params)
    sim.init_waterline_map(yw)
    b = lbm.make_sphere(n, n, n, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"])
    b.vx, b.vy, b.vz = sph["vx"], sph["vy"], sph["vz"]
    return sim, b, params, yw, sph


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(n, seed, steps, parallel=False):
    """The oracle on the same scene: MLUPS over `steps` steps, on one thread, or
    with parallel=True the OpenMP build (same results) on OMP_NUM_THREADS
    threads (default: every core)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py as O
    bits = "32_omp" if parallel else 32
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() if parallel else 1
    params, yw, sph = scenegen.config1(n, n, n, seed)
    s = O.OSim(n, n, n, bits=bits)
    s.set_params(**params)
    s.init_waterline_map(yw)
    b = O.sphere(n, n, n, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"], bits=bits)
    b.vx, b.vy, b.vz = sph["vx"], sph["vy"], sph["vz"]
    s.coupled(b, 1)  # first step includes the initial dead-cell sweep
    t0 = time.perf_counter()
    s.coupled(b, steps)
    dt = time.perf_counter() - t0
    how = (f"the OpenMP build (z-planes of each per-cell pass in parallel) on {threads} threads"
           if parallel else "1 thread")
    return {"value": round(n ** 3 * steps / dt / 1e6, 3), "unit": "MLUPS", "cores": threads,
            "kind": "port",
            "sample": f"{n}^3 config-1 scene, {steps} coupled steps after 1 untimed, oracle "
                      f"(C restatement of sim.js/solid.js, FP64 arithmetic), {how} of "
                      f"{cpu_model()}"}


def batch_run(args, storage, steps):
    """Config 2: the seeded scenes stepped together by the same launches
    (lbm_batch) with populations stored as `storage`: warm-up, the launch-shape
    autotuner, then `steps` timed steps; aggregate MLUPS and per-kernel times."""
    scenes = scenegen.config2(args.scenes, seed=args.seed + 1)
    batch = lbm.Batch()
    batch.set_stream_split(args.batch_split)
    cells = 0
    keep = []
    for (nx, ny, nz), params, yw, sph in scenes:
        sim = lbm.Sim(nx, ny, nz)
        sim.set_params(**params)
        sim.init_waterline_map(yw)
        sim.set_storage(storage)
        b = lbm.make_sphere(nx, ny, nz, sph["cx"], sph["cy"], sph["cz"], sph["r"], sph["density"])
        batch.add(sim, b)
        keep.append((sim, b))
        cells += nx * ny * nz
    print(f"[bench] batch {storage}: {len(scenes)} scenes, {cells} cells; warmup {args.warmup}",
          file=sys.stderr, flush=True)
    batch.step(args.warmup)
    # Launch shapes from the autotuner (its trial steps are ordinary steps).
    tune = batch.autotune(args.tune_steps) if args.tune_steps > 0 else None
    # The timed call launches as in production (steady steps replayed from a
    # captured graph); lbm_batch_step returns after its stream has drained.
    t0 = time.perf_counter()
    batch.step(steps)
    dt = time.perf_counter() - t0
    # Per-kernel times from a separate call with HIP events around every launch
    # (kernels launched one by one).
    ksteps = min(steps, 100)
    batch.set_timing(True)
    batch.step(ksteps)
    kt = {k: batch.kernel_time(k) for k in ("stream_collide", "phase_stream", "stream_macro",
                                           "body", "band", "collide_all")}
    batch.set_timing(False)
    mlups = cells * steps / dt / 1e6
    ms_sc, cnt_sc = kt["stream_collide"]
    avg_s = ms_sc / max(cnt_sc, 1) / 1e3
    kb = KERNEL_BYTES_PER_LU[storage]
    achieved = kb * cells / avg_s / 1e9 if cnt_sc else None
    finite = all(bool(np.isfinite(s.get("phi")).all()) for s, _ in keep[:4])
    batch.close()
    for sim, _ in keep:
        sim.close()
    return {
        "storage": storage, "mlups": round(mlups, 2), "ms_per_step": round(dt / steps * 1e3, 4),
        "steps": steps, "scenes": len(scenes), "cells": cells,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": None, "kernel": STREAM_KERNEL[storage].replace("k1_", "km_")
                     + " (all scenes per launch)",
                     "bytes_per_lu": kb, "avg_ms": round(avg_s * 1e3, 4),
                     "timed_over": f"{ksteps} steps with HIP events after the timed call"},
        "whole_step_hbm_frac": round(STEP_BYTES_PER_LU[storage] * mlups * 1e6 /
                                     (HBM_PEAK_GBS * 1e9), 4),
        "kernel_ms": {k: {"total": round(v[0], 3), "launches": v[1]} for k, v in kt.items()},
        "finite": finite,
        "autotune": tune,
    }


BATCH_WORKLOAD = "config 2: {n} seeded scenes, axes 32..96, one sphere each, one batch"


def bench_batch(args, copy_gbps):
    """The batched workload alone (--workload batch): one JSON line."""
    r = batch_run(args, args.storage, args.steps)
    out = {
        "metric": "batched scenes aggregate MLUPS (config 2)",
        "value": r["mlups"], "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ARITH[args.storage], "data": "synthetic",
        "config": {"workload": BATCH_WORKLOAD.format(n=r["scenes"])
                   + f", {STORAGE_NAME[args.storage]} populations",
                   "scenes": r["scenes"], "cells": r["cells"], "storage": args.storage,
                   "arithmetic": ARITH[args.storage], "seed": args.seed + 1},
        "roofline": r["roofline"],
        "whole_step_hbm_frac": r["whole_step_hbm_frac"],
        "copy_probe_gbps": round(copy_gbps, 1),
        "kernel_ms": r["kernel_ms"],
        "finite": r["finite"],
        "autotune": r["autotune"],
    }
    print(json.dumps(out), flush=True)
    return 0


def serve_consume(srv, args, t0, lat, first):
    """Reads frames until the run's duration ends; returns (stats, seconds)."""
    last_print = t0
    while True:
        now = time.perf_counter()
        if now - t0 >= args.duration:
            break
        fr = srv.next_frame(timeout_ms=100, want_rg8=True, copy=False)
        if now - last_print > 10:
            st = srv.stats()
            print(f"[bench] serve t={now - t0:.0f}s active={st['active']} pending={st['pending']} "
                  f"frames={st['frames']}", file=sys.stderr, flush=True)
            last_print = now
        if fr is None:
            continue
        info, _ = fr
        (first if info.index == 0 else lat).append(info.latency)
    return srv.stats(), time.perf_counter() - t0


def serve_run(args, storage, duration):
    """Served scenes (config 4): Poisson arrivals, admission by memory budget,
    continuous batching, an RG8 frame every 10 steps, for `duration` seconds;
    returns aggregate MLUPS, frame latency and time-to-first-frame percentiles."""
    import threading
    reqs = scenegen.config4(duration, rate=args.rate, seed=args.seed + 3)
    srv = lbm.Server(int(args.budget_gb * (1 << 30)), max_active=args.max_active)
    t0 = time.perf_counter()
    stop = threading.Event()
    submitted = [0]

    def feeder():
        for (ta, (nx, ny, nz), params, yw, sph, steps) in reqs:
            delay = ta - (time.perf_counter() - t0)
            if delay > 0 and stop.wait(delay):
                return
            b = lbm.make_sphere(nx, ny, nz, sph["cx"], sph["cy"], sph["cz"], sph["r"],
                                sph["density"])
            try:
                srv.submit(nx, ny, nz, params, steps, frame_every=10, waterline=yw, sphere=b,
                           storage=storage)
                submitted[0] += 1
            except lbm.LbmError:
                pass

    th = threading.Thread(target=feeder, daemon=True)
    th.start()
    lat, first = [], []
    sargs = argparse.Namespace(duration=duration)
    try:
        st, dt = serve_consume(srv, sargs, t0, lat, first)
    finally:
        stop.set()
        th.join()
        srv.close()
    mlups = st["lattice_updates"] / dt / 1e6
    pct = lambda v, p: round(float(np.percentile(v, p)) * 1e3, 2) if v else None
    return {
        "storage": storage, "mlups": round(mlups, 2), "duration_s": round(dt, 1),
        "frame_latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "count": len(lat)},
        "first_frame_latency_ms": {"p50": pct(first, 50), "p99": pct(first, 99),
                                   "count": len(first)},
        "requests": {"offered": len(reqs), "submitted": submitted[0], "admitted": st["admitted"],
                     "completed": st["completed"], "max_active": st["max_active_seen"]},
        "gpu_busy_frac": round(st["busy_seconds"] / dt, 4),
        "frames": len(lat) + len(first),
    }


SERVE_WORKLOAD = ("config 4: Poisson arrivals at 20/s, axes log-uniform 32..192, step budgets "
                  "500..5000, RG8 frame every 10 steps, admission by memory budget, continuous "
                  "batching")


def bench_serve(args):
    """The serving workload alone (--workload serve): one JSON line."""
    r = serve_run(args, args.storage, args.duration)
    out = {
        "metric": "served scenes aggregate MLUPS, p99 frame latency (config 4)",
        "value": r["mlups"], "unit": "MLUPS", "n_gpus": 1, "steps": None,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": ARITH[args.storage], "data": "synthetic",
        "config": {"workload": SERVE_WORKLOAD, "rate_per_s": args.rate,
                   "duration_s": args.duration, "budget_gib": args.budget_gb,
                   "max_active": args.max_active, "storage": args.storage,
                   "seed": args.seed + 3},
        "serve": r,
    }
    print(json.dumps(out), flush=True)
    return 0


def single_leg(args, storage, copy_gbps):
    """The config-1 scene on one GPU with populations stored as `storage`:
    warm-up, the launch-shape autotuner, then the timed steps.

    The timed steps continue a live run (lbm_step_live): the warm-up ends live,
    the timed call continues it and ends live, so neither the per-call prologue
    (retag, fresh pass and collide of every cell) nor the stream-only step that
    puts a call's end state in the reference's form falls inside the timed
    region; every timed step does the whole coupled step.  The same steps in one
    call that starts and ends in the reference's state are timed as well
    (`reference_state_call`).  Kernel times come from a further call with HIP
    events around every launch on the library's stream."""
    n = args.size
    pattern_gbps = lbm.pull_bandwidth(n, n, n, storage, 10)
    sim, b, params, yw, sph = build_scene(n, args.seed)
    sim.set_storage(storage)
    print(f"[bench] {storage}: scene {n}^3 built; autotune, warmup {args.warmup}",
          file=sys.stderr, flush=True)
    # Launch shapes from the autotuner (its trial steps are ordinary steps).
    tune = sim.autotune(b, args.tune_steps) if args.tune_steps > 0 else None
    # The warm-up is long enough to capture the chosen launch graph.
    sim.coupled_live(b, max(args.warmup, (tune["graph_steps"] + 3) if tune else 1))
    print(f"[bench] {storage}: timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sim.coupled_live(b, args.steps)
    dt = time.perf_counter() - t0
    # Kernel durations: the next steps with events around every launch (one by
    # one; the call starts from the live state and ends in the reference's).
    sim.set_timing(True)
    t1 = time.perf_counter()
    sim.coupled(b, args.steps)
    dt_ev = time.perf_counter() - t1
    kt = {k: sim.kernel_time(k) for k in ("stream_collide", "phase_stream", "stream_macro",
                                          "body", "band", "collide_all")}
    sim.set_timing(False)
    # One call of the same length from and to the reference's state.
    t2 = time.perf_counter()
    sim.coupled(b, args.steps)
    dt_ref = time.perf_counter() - t2
    cells = n ** 3
    mlups = cells * args.steps / dt / 1e6
    ms_sc, cnt_sc = kt["stream_collide"]
    avg_s = ms_sc / max(cnt_sc, 1) / 1e3
    # The row march runs for float32 rows of at most 256 cells when the
    # autotuner (or the default) chose it; its launch pair replaces the phase
    # pass and the stream kernel.
    split = tune["stream_split"] if tune else 2
    march = storage == "f32" and split == 2 and n <= 256
    kb = MARCH_BYTES_PER_LU if march else KERNEL_BYTES_PER_LU[storage]
    kname = MARCH_KERNEL if march else STREAM_KERNEL[storage]
    achieved = kb * cells / avg_s / 1e9 if cnt_sc else None
    phi = sim.get("phi")
    finite = bool(np.isfinite(phi).all())
    sim.close()
    return {
        "storage": storage, "mlups": round(mlups, 2), "ms_per_step": round(dt / args.steps * 1e3, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": measured_traffic(n, storage),
                     "kernel": kname, "bytes_per_lu": kb,
                     "avg_ms": round(avg_s * 1e3, 4),
                     "timed_over": f"{args.steps} steps with HIP events after the timed call "
                                   f"({dt_ev / args.steps * 1e3:.4f} ms/step with events)"},
        "whole_step_hbm_frac": round(STEP_BYTES_PER_LU[storage] * mlups * 1e6 /
                                     (HBM_PEAK_GBS * 1e9), 4),
        "pattern_probe": {"gbps": round(pattern_gbps, 1),
                          "bytes_per_lu": PATTERN_BYTES_PER_LU[storage],
                          "kernel_frac": round(achieved / pattern_gbps, 4) if achieved else None},
        "reference_state_call": {"ms_per_step": round(dt_ref / args.steps * 1e3, 4),
                                 "mlups": round(cells * args.steps / dt_ref / 1e6, 2),
                                 "note": f"one lbm_step call of {args.steps} steps, with the "
                                         "per-call prologue and closing stream-only step"},
        "kernel_ms": {k: {"total": round(v[0], 3), "launches": v[1]} for k, v in kt.items()},
        "finite": finite,
        "sphere": {"cx": b.cx, "cy": b.cy, "cz": b.cz, "vx": b.vx, "vy": b.vy, "vz": b.vz},
        "autotune": tune,
        "sph_r": sph["r"], "params": params,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-par-steps", type=int, default=10,
                    help="steps of the multi-threaded oracle baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-only", action="store_true",
                    help="only run the plain copy kernel (counter calibration)")
    ap.add_argument("--workload", choices=["single", "batch", "serve"], default="single",
                    help="single: the 256^3 scene (headline); batch: 64 seeded scenes of "
                         "32..96 cells per axis stepped together; serve: the request loop "
                         "under Poisson arrivals")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: seconds")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: arrivals per second")
    ap.add_argument("--budget-gb", type=float, default=160.0, help="serve: GiB of scene state")
    ap.add_argument("--max-active", type=int, default=48, help="serve: active scene cap")
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--batch-split", type=int, default=-1, choices=[-1, 0, 1],
                    help="batch: bulk/band-row split of the stream kernel (-1 by storage; "
                         "the autotuner may change it)")
    ap.add_argument("--tune-steps", type=int, default=10,
                    help="single and batch: autotuner trial steps per candidate (0: no "
                         "autotuning)")
    ap.add_argument("--legs", default="f16,batch,serve",
                    help="single workload: extra legs in the line (f16: the FP16-deviation form "
                         "on the same scene; batch: config 2 in both forms; serve: the serving "
                         "loop in both forms)")
    ap.add_argument("--batch-steps", type=int, default=200,
                    help="single workload: timed steps of each batch leg")
    ap.add_argument("--serve-seconds", type=float, default=30.0,
                    help="single workload: seconds of each serving leg")
    ap.add_argument("--storage", choices=["f32", "f16", "bf16"], default="f32",
                    help="population storage: f32 (reference-exact, FP64 arithmetic) or a "
                         "16-bit deviation (FP32 arithmetic)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the coupled step on a single MI355X "
              "(multi-GPU sharding is a design note only).", file=sys.stderr)
        return 2
    copy_gbps = lbm.copy_bandwidth(1 << 31, 10)
    if args.copy_probe_only:
        print(json.dumps({"copy_probe_gbps": round(copy_gbps, 1), "bytes": 1 << 31, "reps": 10}))
        return 0
    if args.workload == "batch":
        return bench_batch(args, copy_gbps)
    if args.workload == "serve":
        return bench_serve(args)
    n = args.size
    head = single_leg(args, args.storage, copy_gbps)
    out = {
        "metric": METRIC,
        "value": head["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ARITH[args.storage],
        "data": "synthetic",
        "config": {"workload": f"config 1: {n}^3 single scene, coupled step (D3Q27 hydro + D3Q7 "
                               f"phase + sphere), {STORAGE_NAME[args.storage]} populations",
                   "nx": n, "ny": n, "nz": n, "storage": args.storage,
                   "arithmetic": ARITH[args.storage], "seed": args.seed,
                   "sphere_r": head["sph_r"], "params": head["params"],
                   "timed": "steps of a continuing run (lbm_step_live), inputs resident in HBM"},
        "roofline": head["roofline"],
        "whole_step_hbm_frac": head["whole_step_hbm_frac"],
        "copy_probe_gbps": round(copy_gbps, 1),
        # The stream kernel's population pattern alone (34 shifted directions
        # pulled and stored, no arithmetic): the ceiling its memory pattern allows.
        "pattern_probe": head["pattern_probe"],
        "reference_state_call": head["reference_state_call"],
        "kernel_ms": head["kernel_ms"],
        "finite": head["finite"],
        "sphere": head["sphere"],
        "autotune": head["autotune"],
    }
    # The compressed form (FP16 deviations) on the same scene and the served
    # workload in both forms: the metric's ratio and frame latency.
    if args.storage == "f32" and "f16" in args.legs:
        f16 = single_leg(args, "f16", copy_gbps)
        for k in ("sph_r", "params", "sphere"):
            f16.pop(k)
        f16["over_f32"] = round(f16["mlups"] / head["mlups"], 3)
        out["f16"] = f16
    if "batch" in args.legs and args.batch_steps > 0:
        batch = {}
        for st in ("f32", "f16"):
            batch[st] = batch_run(args, st, args.batch_steps)
        batch["f16_over_f32"] = round(batch["f16"]["mlups"] / batch["f32"]["mlups"], 3)
        batch["workload"] = BATCH_WORKLOAD.format(n=batch["f32"]["scenes"])
        out["batch"] = batch
    if "serve" in args.legs and args.serve_seconds > 0:
        serve = {}
        for st in ("f32", "f16"):
            print(f"[bench] serving, {st}, {args.serve_seconds:.0f} s", file=sys.stderr,
                  flush=True)
            serve[st] = serve_run(args, st, args.serve_seconds)
        serve["f16_over_f32"] = round(serve["f16"]["mlups"] / serve["f32"]["mlups"], 3)
        serve["workload"] = (SERVE_WORKLOAD + f"; {args.serve_seconds:.0f} s per storage form "
                             "(the configuration's 10 minutes do not fit the benchmark window; "
                             "--workload serve --duration 600 runs them)")
        out["serve"] = serve
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(n, args.seed, args.cpu_steps)
        print("[bench] cpu baseline, all cores", file=sys.stderr, flush=True)
        out["cpu_baseline_parallel"] = cpu_baseline(n, args.seed, args.cpu_par_steps, True)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
