#!/usr/bin/env python3
"""Benchmark of the MI355X SPH step engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1m|8m|16k|batch256|batch256-plain|batch64]

One "step" is SimulationEngine::Update(dt, 1) over the whole workload: density,
forces + integration + rehash, scan and scatter. The default workload is the
single-scene dam break of 1,048,576 particles (128x64x128 jittered lattice at
spacing h/2 in a 12.8x6.4x12.8 wall-bounded box, memory order shuffled, dt
1e-3). Inputs are resident in HBM when the timed region starts. Prints one JSON
line with particle-updates/s, the roofline of the dominant kernel (HIP events on
the engine's stream over the timed region) and the CPU baseline (the restated
reference CPU engine in oracle/, timed on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "particle-updates/s at 1M & 8M; batched particle-updates/s; % of HBM peak"

CONFIGS = {
    # name: (scene factory, dt, default steps, options the workload states)
    "16k": ("cpu_16k", 1e-3, 200, {}),
    "1m": ("gpu_1m", 1e-3, 1000, {}),
    "8m": ("gpu_8m", 5e-4, 500, {}),
    # 256 scenes advance together, each step replayed from one captured hipGraph.
    "batch256": ("batch_256", 1e-3, 500, {"graph": True}),
    # The same scenes, one engine (and one launch per stage) per scene, over the
    # same 500-step window.
    "batch256-plain": ("batch_256", 1e-3, 500, {}),
    # 64 scenes in mixed precision, neighbour positions stored as FP16 offsets
    # from their cell corners.
    "batch64": ("batch_64_mixed", 1e-3, 200, {"graph": True, "mixed": True, "stored": True}),
    # The scene the reference's application loads (triangleTestScene.json with its
    # 1,096-triangle mesh), stepped as the application steps it.
    "triangle": ("triangle_test_scene", 1e-2, 200, {}),
}


def stage_bytes(n: int, M: int, Ms: int, L: int):
    """Algorithmic HBM bytes per launch of each stage (FP32 state, each compulsory
    read or write counted once; neighbour gathers excluded)."""
    return {
        # pos r 12, p w 4; start table r; static table r
        "density": 16 * n + 4 * (M + 1) + 4 * (Ms + 1),
        # pos r 12, vel r 12, p r 4, pos w 12, vel w 12, key w 4; starts r + counts w; static r
        "force": 56 * n + 8 * (M + 1) + 4 * (Ms + 1),
        # counts r, starts w, next histogram cleared
        "scan": 12 * L,
        # key r 4, pos r/w 24, vel r/w 24, id r/w 8; cursors
        "scatter": 60 * n + 4 * (M + 1),
    }


def step_bytes(n: int, M: int, Ms: int) -> int:
    """B_alg per step of the measurement plan: 144 N + 24 (M + 1) + 8 (M_s + 1)."""
    return 144 * n + 24 * (M + 1) + 8 * (Ms + 1)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# Kernel of each timed stage, per neighbour-pass variant.
KERNELS = {
    "cellwave": {"density": "k_density_cw", "force": "k_force_cw"},
    "runs": {"density": "k_density_runs", "force": "k_force_runs"},
    "tile": {"density": "k_density_tile", "force": "k_force_tile"},
    "cells": {"density": "k_density", "force": "k_force"},
}


def _profile_value(kernel: str, workload: str, steps: int, kinds, key: str):
    """Value of the line '<key>: X' of a committed rocprofv3 summary, and its file:
    the summary of this step window (profiles/<kernel>_<workload>-<steps>-steps_<kind>.md)
    if there is one, else the workload's (profiles/<kernel>_<workload>_<kind>.md).
    The density pass's work per step grows along a run, so the window matters."""
    for name in (f"{workload}-{steps}-steps", workload):
        for kind in kinds:
            rel = os.path.join("profiles", f"{kernel}_{name}_{kind}.md")
            path = os.path.join(ROOT, rel)
            if not os.path.exists(path):
                continue
            for line in open(path):
                if line.startswith(key + ":"):
                    try:
                        return float(line.split(":", 1)[1]), rel
                    except ValueError:
                        return None, rel
    return None, None


def load_traffic(kernel: str, workload: str, steps: int):
    """HBM bytes per launch of a kernel (FETCH_SIZE + WRITE_SIZE, calibrated) from
    a committed --pmc or trace summary (line 'traffic_bytes: X'), if any."""
    return _profile_value(kernel, workload, steps, ("pmc", "trace"), "traffic_bytes")[0]


def load_valu_busy(kernel: str, workload: str, steps: int):
    """Fraction of SIMD cycles issuing VALU instructions (SQ_INSTS_VALU x 4 over
    1024 SIMDs x kernel cycles) from a committed SQ counter summary (line
    'valu_busy: X'), and that summary's path, if any."""
    return _profile_value(kernel, workload, steps, ("sq",), "valu_busy")


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def baseline_threads() -> int:
    """Threads for the CPU baseline: the CPUs this process may run on (the
    reference sizes its pool by std::thread::hardware_concurrency), capped by the
    job's CPU share when the host states one in OMP_NUM_THREADS."""
    try:
        t = len(os.sched_getaffinity(0))
    except AttributeError:
        t = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        t = min(t, int(share))
    return max(1, t)


def time_oracle(orc, scene, dt: float, budget_s: float, threads: int):
    """Particle-updates/s of the oracle on `scene` with `threads` threads: one
    warm-up Update, then as many steps as fit the budget (2 .. 200)."""
    ref = orc.Oracle(scene.params, scene.pos, scene.vel, scene.static_pos, threads=threads)
    t0 = time.perf_counter()
    ref.update(dt, 1)
    t1 = time.perf_counter() - t0
    steps = max(2, min(200, int(budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    ref.update(dt, steps)
    el = time.perf_counter() - t0
    ref.close()
    return scene.n * steps / el, steps, el


def cpu_baseline(scene, dt: float, budget_s: float, threads: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    value, steps, el = time_oracle(orc, scene, dt, budget_s, threads)
    one, steps1, el1 = time_oracle(orc, scene, dt, budget_s / 2, 1) if threads > 1 else (value, steps, el)
    return {
        "value": value,
        "unit": "particle-updates/s",
        "cores": threads,
        "threads": threads,
        "value_1_thread": one,
        "cpu_model": cpu_model(),
        "host_cpus": os.cpu_count(),
        "kind": "port",
        "sample": f"{scene.name}: {steps} steps of {scene.n} particles after 1 warm-up step, "
                  f"restated reference CPU engine (AoS, particleMap, serial scan, validators off: a stated "
                  f"deviation, the reference's debug validators run every step), {threads} threads, {el:.1f} s; "
                  f"1 thread: {steps1} steps, {el1:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="1m", choices=sorted(CONFIGS))
    ap.add_argument("--graph", action="store_true", help="replay steps from hipGraphs")
    ap.add_argument("--walk", default="cellwave", choices=["cellwave", "tile", "runs", "cells"],
                    help="neighbour passes: cell-group waves (default), LDS-staged bucket tiles, "
                         "HBM Morton runs, or cell by cell")
    ap.add_argument("--neighbours", choices=["auto", "lists", "masks"], default="auto",
                    help="cell-wave walk: how the density pass hands neighbours to the force pass "
                         "(auto: lists up to 1.5M particles, masks above)")
    ap.add_argument("--mixed", action="store_true",
                    help="mixed precision: FP16-offset candidate filter, Kahan-compensated neighbour sums")
    ap.add_argument("--stored", action="store_true",
                    help="mixed precision with the FP16 offsets stored per particle (implies --mixed)")
    ap.add_argument("--on-the-fly", action="store_true",
                    help="mixed precision with the FP16 offsets formed from the positions (overrides a config's --stored)")
    ap.add_argument("--autotune", action="store_true",
                    help="after initialize, let the engine time every walk (and the default walk's launch "
                         "shape) and keep the fastest")
    ap.add_argument("--shape", default=None, metavar="D,F",
                    help="launch shape of the default walk: waves per workgroup of the density and the force "
                         "pass (1, 2, 4 or 8 each)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of CPU baseline work with all threads (half that again with one thread)")
    ap.add_argument("--profile-all", action="store_true", help="time every stage with events (default: dominant)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (HBM counter calibration) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not an implementation)", file=sys.stderr)
        sys.exit(2)


    from sphfluidsimulation_amd import SimulationEngineHIP, scenes
    from sphfluidsimulation_amd._lib import (SPH_FLAG_CELL_WALK, SPH_FLAG_CW_LISTS, SPH_FLAG_CW_MASKS,
                                             SPH_FLAG_FP16_OFFSETS, SPH_FLAG_FP16_STORED, SPH_FLAG_HIP_GRAPH,
                                             SPH_FLAG_RUN_WALK,
                                             SPH_FLAG_TILE_WALK)

    if args.copy_probe:
        e = SimulationEngineHIP(0)
        e.copy_probe(args.copy_probe, 3)
        e.close()
        log(f"[bench] copy probe: 3 copies of {args.copy_probe} bytes")
        return

    factory, dt, default_steps, stated = CONFIGS[args.config]
    args.graph = args.graph or stated.get("graph", False)
    args.stored = (args.stored or stated.get("stored", False)) and not args.on_the_fly
    args.mixed = args.mixed or args.stored or stated.get("mixed", False)
    steps = args.steps if args.steps is not None else default_steps
    t0 = time.perf_counter()
    if factory == "triangle_test_scene":
        from sphfluidsimulation_amd import scene_json
        built = scene_json.load_scene(os.path.join(ROOT, "tests", "golden", "app", "assets", "simulationScenes",
                                                   "triangleTestScene.json"))
    else:
        built = getattr(scenes, factory)()
    scene_list = built if isinstance(built, list) else [built]
    log(f"[bench] {args.config}: {len(scene_list)} scene(s), "
        f"{sum(s.n for s in scene_list)} particles, generated in {time.perf_counter() - t0:.1f} s")

    flags = ((SPH_FLAG_HIP_GRAPH if args.graph else 0) | (SPH_FLAG_CELL_WALK if args.walk == "cells" else 0)
             | (SPH_FLAG_RUN_WALK if args.walk == "runs" else 0) | (SPH_FLAG_TILE_WALK if args.walk == "tile" else 0)
             | (SPH_FLAG_FP16_OFFSETS if args.mixed else 0) | (SPH_FLAG_FP16_STORED if args.stored else 0)
             | (SPH_FLAG_CW_LISTS if args.neighbours == "lists" else 0)
             | (SPH_FLAG_CW_MASKS if args.neighbours == "masks" else 0))
    plain = args.config.endswith("-plain")
    engines = []
    if plain:
        for sc in scene_list:
            e = SimulationEngineHIP(0)
            e.Initialize(sc, flags=flags)
            engines.append(e)
    else:
        e = SimulationEngineHIP(0)
        e.Initialize(scene_list if len(scene_list) > 1 else scene_list[0], flags=flags)
        engines.append(e)
    if args.shape:
        d, f = (int(x) for x in args.shape.split(","))
        for e in engines:
            e.set_launch_shape(d, f)
    if args.autotune:
        chosen = {e.autotune(dt, 4) for e in engines}
        args.walk = "/".join(sorted(chosen))
        handoff = {e.neighbours for e in engines if e.neighbours}
        if handoff:
            args.neighbours = "/".join(sorted(handoff))
        log(f"[bench] autotune chose {args.walk} ({args.neighbours})")
    shapes = sorted({e.launch_shape() for e in engines})
    n_total = sum(e.n for e in engines)
    M_total = sum(e.hash_map_size(s) for e in engines for s in range(e.n_scenes))
    Ms_total = sum(sc.n_static for sc in scene_list)
    L_total = sum(e.hash_map_size(s) + 1 for e in engines for s in range(e.n_scenes))
    log(f"[bench] initialized: N={n_total} M={M_total} walls={Ms_total}")

    def run(k):
        for _ in range(k):
            for e in engines:
                e.Update(dt, 1)

    run(args.warmup)
    for e in engines:
        e.synchronize()
    log(f"[bench] warm-up {args.warmup} steps done")

    # Timed region: events around the stages we report (the engine's own stream).
    # hipGraph replay carries no per-kernel events, so with --graph the timed
    # region replays graphs untouched and the stage times come from an eager
    # window of the same length (at most 100 steps) right after it. With one
    # engine per scene the events would add host work per engine to the timed
    # region, so those runs take their stage times from such a window too.
    prof_stages = True if args.profile_all else ["force", "density"]
    events_after = args.graph or plain

    def collect():
        out = {}
        for e in engines:
            for k, (ms, cnt) in e.stage_times().items():
                a = out.setdefault(k, [0.0, 0])
                a[0] += ms
                a[1] += cnt
            e.set_profiling(None)
        return out

    # In the timed region each event record costs the stream a few microseconds,
    # so the two stages are bracketed on alternating steps (SPH_PROFILE_ROTATE):
    # each is timed over the whole region, on every other step.
    rotate = not args.profile_all
    if not events_after:
        for e in engines:
            e.set_profiling(prof_stages, rotate=rotate)
    t0 = time.perf_counter()
    run(steps)
    for e in engines:
        e.synchronize()
    elapsed = time.perf_counter() - t0
    if events_after:
        for e in engines:
            e.set_profiling(prof_stages)
        run(min(steps, 100))
        for e in engines:
            e.synchronize()
    times = collect()
    log(f"[bench] timed {steps} steps in {elapsed:.3f} s; stage ms: "
        + ", ".join(f"{k}={v[0]:.1f}/{v[1]}" for k, v in times.items()))

    pu_s = n_total * steps / elapsed
    ms_per_step = elapsed * 1e3 / steps
    sb = stage_bytes(n_total, M_total, Ms_total, L_total)
    # Dominant kernel: the timed stage with the largest total time.
    timed = {k: v for k, v in times.items() if v[1] > 0}
    dom = max(timed, key=lambda k: timed[k][0])
    avg_s = timed[dom][0] / 1e3 / (timed[dom][1] / len(engines))
    achieved = sb[dom] / avg_s / 1e9
    workload = scene_list[0].name if len(scene_list) == 1 else f"{args.config}-{len(scene_list)}-scenes"
    walk_kernels = KERNELS.get(args.walk, {})
    kernel = walk_kernels.get(dom, "k_" + dom)
    traffic = load_traffic(kernel, workload, steps)
    valu_busy, valu_src = load_valu_busy(kernel, workload, steps)
    b_step = step_bytes(n_total, M_total, Ms_total)
    result = {
        "metric": METRIC,
        "value": pu_s,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "particles": n_total,
            "static_particles": Ms_total,
            "scenes": len(scene_list),
            "launch": "one engine per scene" if plain else "batched" if len(scene_list) > 1 else "single scene",
            "dt": dt,
            "hip_graph": bool(args.graph),
            "stage_times_from": "eager window after the timed region" if events_after
            else "timed region, density and force on alternating steps" if rotate else "timed region",
            "walk": args.walk,
            "neighbours": args.neighbours,
            "autotune": bool(args.autotune),
            "launch_shape": "/".join(f"density {d}w, force {f}w" for d, f in shapes),
            "precision": ("mixed (fp16 offsets stored per particle, kahan sums)" if args.stored
                          else "mixed (fp16 filter, kahan sums)" if args.mixed else "fp32"),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "avg_launch_us": avg_s * 1e6,
            "algorithmic_bytes_per_launch": sb[dom],
        },
        # What does bound the dominant kernel when it is not HBM: its vector ALU
        # issue, from the committed counter summary of this window (null if none).
        "issue_bound": {"unit": "VALU issue", "busy": valu_busy, "source": valu_src},
        "step_roofline": {
            "b_alg_per_step": b_step,
            "achieved_GBs": b_step * steps / elapsed / 1e9,
            "frac": b_step * steps / elapsed / 1e9 / HBM_PEAK_GBS,
        },
        "stage_ms_per_step": {k: v[0] / max(1, v[1] / len(engines)) for k, v in timed.items()},
        # Achieved GB/s of each timed stage: its algorithmic bytes per launch over its average launch time.
        "stage_GBs": {k: sb[k] / (v[0] / 1e3 / max(1, v[1] / len(engines))) / 1e9 for k, v in timed.items()},
    }
    for e in engines:
        e.close()

    if not args.no_cpu_baseline and len(scene_list) == 1:
        threads = baseline_threads()
        log(f"[bench] CPU baseline: {threads} threads (host: {os.cpu_count()} CPUs), then 1 thread")
        result["cpu_baseline"] = cpu_baseline(scene_list[0], dt, args.cpu_budget, threads)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
