#!/usr/bin/env python3
"""Benchmark: D3Q19 MLUPS and fraction of the HBM roofline on one MI355X.

Workload (default): D3Q19 BGK channel flow past a sphere, 256x128x128 FP32,
sphere r=16 at (64,64,64), equilibrium inlet u=0.05, zero-gradient outflow,
periodic y, z periodic through the z-slab halo planes (pack fused into the
sweep, device copy, unpack fused into the next sweep), tau=0.56.  A step is
one time step of the whole lattice.  Inputs are resident in HBM before the
timed region; the timed region is exactly --steps steps between two device
synchronisations.

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lattibol_0.3.31f_amd"))

PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16", "f64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=0, help="oracle steps for the CPU baseline (0 = auto)")
    ap.add_argument("--kernel-launches", type=int, default=200)
    return ap.parse_args()


def traffic_from_profiles(storage):
    """HBM bytes per sweep measured with rocprofv3 --pmc (committed summary)."""
    p = os.path.join(ROOT, "profiles", "d3q19_sphere_256x128x128_%s_traffic.json" % storage)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as fh:
            return json.load(fh).get("hbm_bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(cfg, solid, init, steps_hint):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import numpy as np
    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    orc = O.PullOracle(cfg, solid)
    orc.init_equilibrium(init["rho"], init["ux"], init["uy"], init["uz"])
    orc.run(1)   # warm
    steps = steps_hint
    if steps <= 0:
        t0 = time.perf_counter()
        orc.run(2)
        per = (time.perf_counter() - t0) / 2
        steps = max(2, min(200, int(15.0 / max(per, 1e-6))))
    t0 = time.perf_counter()
    orc.run(steps)
    dt = time.perf_counter() - t0
    n = cfg["nx"] * cfg["ny"] * cfg["nz"]
    return {"value": round(n * steps / dt / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": "%d steps of the full %dx%dx%d lattice, FP32 oracle (C++ host loop, OpenMP over z-planes, "
                      "same arithmetic as the GPU path), %.1f s" % (steps, cfg["nx"], cfg["ny"], cfg["nz"], dt)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported "
              "(the z-slab halo exchange is exercised on one device)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import lattibol_amd as lb
    from lattibol_amd import scenes

    cfg, solid, init = scenes.sphere_channel(storage=args.storage)
    nodes = cfg["nx"] * cfg["ny"] * cfg["nz"]
    esz = {"f32": 4, "f16": 2, "bf16": 2, "f64": 8}[args.storage]
    bytes_per_update = 2 * 19 * esz
    print("bench: creating solver %dx%dx%d %s" % (cfg["nx"], cfg["ny"], cfg["nz"], args.storage), file=sys.stderr,
          flush=True)
    sol = lb.Solver(cfg, solid)
    sol.init_equilibrium(init["rho"], init["ux"], init["uy"], init["uz"])
    launch = sol.launch_config
    sol.step(args.warmup)
    sol.sync()
    t0 = time.perf_counter()
    sol.step(args.steps)
    sol.sync()
    dt = time.perf_counter() - t0
    mlups = nodes * args.steps / dt / 1e6
    # sanity: the run is finite and mass is positive
    tot = sol.reduce()
    if not np.isfinite(tot).all() or tot[0] <= 0:
        raise SystemExit("bench: non-finite state after the timed run")
    print("bench: %.1f MLUPS over %d steps (%.3f s)" % (mlups, args.steps, dt), file=sys.stderr, flush=True)

    # dominant kernel: the fused sweep, timed with HIP events on its stream
    ms = sol.time_step_kernel(args.kernel_launches)
    alg_bytes = bytes_per_update * nodes
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(args.storage)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "kernel": "k_step<19,%s> fused pull sweep" % args.storage,
                "kernel_ms": round(ms, 5), "algorithmic_bytes_per_launch": alg_bytes}
    sol.close()

    cpu = None
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(scenes.sphere_channel(storage="f32")[0], solid, init, args.cpu_steps)

    line = {
        "metric": "D3Q19 MLUPS + % HBM peak; time-to-steady-state, warm-start vs plain LBM",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "fp32", "f16": "fp16-storage/fp32-compute", "bf16": "bf16-storage/fp32-compute",
                  "f64": "fp64"}[args.storage],
        "data": "synthetic",
        "config": {"workload": "D3Q19 BGK channel past a sphere 256x128x128 (r=16 at (64,64,64)), inlet u=0.05, "
                               "zero-gradient outflow, periodic y, z-slab halo planes, tau=0.56, %s" % args.storage,
                   "nodes": nodes, "bytes_per_update": bytes_per_update, "launch": launch,
                   "pct_hbm_peak_whole_step": round(100 * mlups * 1e6 * bytes_per_update / (PEAK_HBM_GBS * 1e9), 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
