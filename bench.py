#!/usr/bin/env python3
"""Benchmark: D3Q19 MLUPS and fraction of the HBM roofline on one MI355X.

Workload (default): D3Q19 BGK channel flow past a sphere, 256x128x128 FP32,
sphere r=16 at (64,64,64), equilibrium inlet u=0.05, zero-gradient outflow,
periodic y, z periodic through the z-slab halo planes (pack fused into the
sweep, unpack fused into the next sweep, two alternating halo buffers),
tau=0.56.  A step is
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
    ap.add_argument("--storage", default="f32", choices=["f32", "f32s", "f16", "bf16", "f64"])
    ap.add_argument("--block-x", type=int, default=0, help="fix the sweep's block width (0 = autotune)")
    ap.add_argument("--xcd-remap", type=int, default=-1, help="fix the XCD tile remap (0/1, -1 = autotune)")
    ap.add_argument("--mem-mode", type=int, default=-1,
                    help="fix the load/store mode (0 plain, 1 non-temporal stores, 2 non-temporal loads+stores)")
    ap.add_argument("--layout", type=int, default=-1,
                    help="fix the population layout (0 [i][z][y][x], 1 [z][y][i][x]; -1 = autotune)")
    ap.add_argument("--block-y", type=int, default=0,
                    help="rows per tile of the one-pass Shan-Chen sweeps (0 = autotune)")
    ap.add_argument("--sc-zchunk", type=int, default=0,
                    help="z-planes per workgroup of the one-pass Shan-Chen sweeps (0 = autotune)")
    ap.add_argument("--sc-sweep", type=int, default=-1,
                    help="Shan-Chen step form (0 density pass + sweep, 1 / 2 one z-marching sweep on tiles, "
                         "3 row-slab one-pass sweep of periodic boxes; -1 = autotune)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--error-steps", type=int, default=5000,
                    help="steps of the FP64/FP32/FP16/BF16 error runs of the sphere workload (0 = skip)")
    ap.add_argument("--cpu-steps", type=int, default=0, help="oracle steps for the CPU baseline (0 = auto)")
    ap.add_argument("--kernel-launches", type=int, default=200)
    ap.add_argument("--workload", default="sphere",
                    choices=["sphere", "cavity", "spinodal", "ensemble", "warmstart", "unet", "reference", "copy"],
                    help="sphere = D3Q19 channel past a sphere (the headline); the others print their own line")
    ap.add_argument("--train-samples", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--heldout", type=int, default=256)
    ap.add_argument("--max-steady-steps", type=int, default=300000)
    ap.add_argument("--spinodal-storages", default="f64,f32,bf16,f16,bf16_nearest,f16_nearest",
                    help="storages of the spinodal workload (comma list; f64 is the reference of the error "
                         "columns; bf16 / f16 round stochastically, the Shan-Chen default; *_nearest round to "
                         "nearest even, for comparison)")
    return ap.parse_args()


def spinodal_traffic(storage):
    """HBM bytes per Shan-Chen step of the tuned form, measured with
    rocprofv3 --pmc on the 192^3 workload (committed summary; the 16-bit
    storages are profiled with nearest-even rounding, which moves the same
    bytes)."""
    p = os.path.join(ROOT, "profiles", "d3q19_spinodal_192_traffic.json")
    try:
        with open(p) as fh:
            return json.load(fh)["hbm_bytes_per_step"][storage]
    except Exception:
        return None


def profile_field(name, key):
    """A figure of a committed rocprofv3 summary under profiles/ (None if absent)."""
    try:
        with open(os.path.join(ROOT, "profiles", name)) as fh:
            return json.load(fh).get(key)
    except Exception:
        return None


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


def cpu_info():
    """Cores of the host and its CPU model (from /proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {"nproc": os.cpu_count(), "cpus_allowed": affinity, "cpu_model": model}


def cpu_threads():
    """Oracle threads: the CPUs this process may use, at most 16 (the GPU
    box's CPU share per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(cfg, solid, init, steps_hint):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import numpy as np
    threads = cpu_threads()
    O.set_threads(threads)
    orc = O.PullOracle(cfg, solid)
    orc.init_equilibrium(init["rho"], init["ux"], init["uy"], init["uz"])
    orc.run(3)   # warm (threads, pages)
    steps = steps_hint
    if steps <= 0:
        t0 = time.perf_counter()
        orc.run(8)
        per = (time.perf_counter() - t0) / 8
        steps = max(2, min(2000, int(15.0 / max(per, 1e-6))))   # about 15 s of CPU work
    t0 = time.perf_counter()
    orc.run(steps)
    dt = time.perf_counter() - t0
    n = cfg["nx"] * cfg["ny"] * cfg["nz"]
    return {"value": round(n * steps / dt / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            **cpu_info(),
            "sample": "%d steps of the full %dx%dx%d lattice, FP32 oracle (C++ host loop, OpenMP over z-planes, "
                      "same arithmetic as the GPU path), %.1f s" % (steps, cfg["nx"], cfg["ny"], cfg["nz"], dt)}


def sphere_errors(steps):
    """C2's storage error: the sphere channel run `steps` steps in FP64, FP32,
    FP16 and BF16 (shifted deviations, nearest-even) from the same start;
    velocity and density of each against the FP64 run (relative L2 over the
    lattice, maximum absolute difference) and the total mass (the channel is
    open, so its mass changes physically: compared with the FP64 run's)."""
    import numpy as np
    import lattibol_amd as lb
    from lattibol_amd import scenes
    ref = None
    out = {}
    for st in ("f64", "f32", "f16", "bf16"):
        cfg, solid, init = scenes.sphere_channel(storage=st)
        sol = lb.Solver(cfg, solid)
        sol.init_equilibrium(init["rho"], init["ux"], init["uy"], init["uz"])
        sol.step(steps)
        mass = sol.reduce()[0]
        m = sol.macros()
        sol.close()
        u = np.stack([m["ux"], m["uy"], m["uz"]]).astype(np.float64)
        rho = m["rho"].astype(np.float64)
        if ref is None:
            ref = (u, rho, mass)
        du, dr = u - ref[0], rho - ref[1]
        out[st] = {"u_l2_rel": float(np.linalg.norm(du) / np.linalg.norm(ref[0])),
                   "u_max_abs": float(np.abs(du).max()),
                   "rho_l2_rel": float(np.linalg.norm(dr) / np.linalg.norm(ref[1])),
                   "rho_max_abs": float(np.abs(dr).max()),
                   "mass_rel_vs_fp64": float((mass - ref[2]) / ref[2])}
        log("C2 error %s after %d steps: %s" % (st, steps, json.dumps(out[st])))
    return {"steps": steps, "u_in": 0.05, "storage": out}


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def common(metric, value, unit, args, steps, ms, dtype, config, higher=True, **extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": args.warmup,
         "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
         "data": "synthetic", "config": config}
    d.update(extra)
    return d


def bench_cavity(args):
    """C1: D2Q9 lid-driven cavity 256^2, tau 0.6, lid u=0.1, 10,000 steps (launch-bound parity config)."""
    import lattibol_amd as lb
    from lattibol_amd import scenes
    cfg, solid, init = scenes.cavity()
    sol = lb.Solver(cfg, solid)
    sol.init_equilibrium(init["rho"], init["ux"], init["uy"])
    sol.step(args.warmup)
    steps = args.steps
    sol.prepare(steps)
    sol.sync()
    t0 = time.perf_counter()
    sol.step(steps)
    sol.sync()
    dt = time.perf_counter() - t0
    n = cfg["nx"] * cfg["ny"]
    tb = sol.launch_config["time_block"]
    ms1 = sol.time_step_kernel(200)          # the single-step sweep, for comparison
    ms = dt / steps * 1e3                    # per step as run (temporally blocked launches in graphs)
    res = sol.reduce()
    return common("D2Q9 MLUPS (lid-driven cavity)", round(n * steps / dt / 1e6, 1), "MLUPS", args, steps,
                  round(ms, 5), "fp32",
                  {"workload": "D2Q9 BGK lid-driven cavity 256x256, tau=0.6, lid u=0.1, half-way bounce-back",
                   "launch": sol.launch_config},
                  roofline={"bound": "hbm", "note": "latency-bound: launch + one HBM round trip per launch",
                            "achieved": round(72 * n / (ms * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS,
                            "unit": "GB/s", "frac": round(72 * n / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                            "traffic": None,
                            "kernel": ("k_step_tb9<%d>, %d steps per launch" % (tb, tb)) if tb else "k_step<9,f32>",
                            "ms_per_step": round(ms, 5), "single_step_kernel_ms": round(ms1, 5)},
                  mass=float(res[0]),
                  cpu_baseline=None if args.no_cpu_baseline else cpu_baseline(cfg, solid, init, args.cpu_steps))


def bench_spinodal(args):
    """C3: D3Q19 Shan-Chen 192^3; FP32 / FP16 / BF16 storage (nearest-even and
    stochastic rounding) against the FP64 run.  Spinodal decomposition is
    chaotic, so after many steps pointwise differences measure a different
    realisation of the pattern; the phase densities (5th / 95th percentiles of
    rho) and the mass drift are the meaningful comparisons."""
    import numpy as np
    import lattibol_amd as lb
    from lattibol_amd import scenes
    steps = args.steps
    out = {}
    fields = {}
    want = set(args.spinodal_storages.split(","))
    runs = [(st, rm) for st, rm in (("f64", "default"), ("f32", "default"), ("bf16", "default"),
                                    ("f16", "default"), ("bf16", "nearest"), ("f16", "nearest"))
            if st + ("_nearest" if rm == "nearest" else "") in want]
    if "f32" not in want:
        raise SystemExit("the spinodal workload reports the f32 run; include f32 in --spinodal-storages")
    for st, rm in runs:
        key = st + ("_nearest" if rm == "nearest" else "")
        cfg, solid, init = scenes.spinodal(storage=st, round_mode=rm, block_x=args.block_x, sc_sweep=args.sc_sweep,
                                           xcd_remap=args.xcd_remap, mem_mode=args.mem_mode, layout=args.layout,
                                           block_y=args.block_y, sc_zchunk=args.sc_zchunk)
        sol = lb.Solver(cfg, solid)
        sol.init_equilibrium(init["rho"], 0.0, 0.0, 0.0)
        m0 = sol.reduce()[0]
        sol.prepare(steps)
        sol.sync()
        t0 = time.perf_counter()
        sol.step(steps)
        sol.sync()
        dt = time.perf_counter() - t0
        m1 = sol.reduce()[0]
        rho = sol.macros()["rho"]
        fields[key] = rho.astype(np.float64)
        # the step's kernels alone, from HIP events on the solver's stream (a
        # graph of back-to-back steps; advances the state, after the fields)
        kms = sol.time_step_kernel(args.kernel_launches)
        n = rho.size
        p5, p95 = np.percentile(rho, [5, 95])
        esz = {"f64": 8, "f32": 4}.get(st, 2)
        out[key] = {"mlups": round(n * steps / dt / 1e6, 1), "seconds": round(dt, 3),
                    "pct_hbm_at_2Q_bytes": round(100 * n * steps / dt * 38 * esz / (PEAK_HBM_GBS * 1e9), 2),
                    "kernel_us_per_step": round(kms * 1e3, 2),
                    "kernel_pct_hbm_at_2Q_bytes": round(100 * n / (kms * 1e-3) * 38 * esz / (PEAK_HBM_GBS * 1e9), 2),
                    "round_mode": sol.launch_config["round_mode"] if esz == 2 else None,
                    "mass_drift_rel": float((m1 - m0) / m0), "rho_min": float(rho.min()), "rho_max": float(rho.max()),
                    "rho_p5": float(p5), "rho_p95": float(p95), "launch": sol.launch_config}
        log("%s: %.1f MLUPS, drift %.2e, rho p5/p95 %.4f %.4f, launch %s" % (
            key, out[key]["mlups"], out[key]["mass_drift_rel"], p5, p95, out[key]["launch"]))
        sol.close()
    ref = fields.get("f64")
    for key in out:
        if key == "f64" or ref is None:
            continue
        d = fields[key] - ref
        out[key]["rho_err_l2_rel_vs_fp64"] = float(np.linalg.norm(d) / np.linalg.norm(ref))
        out[key]["rho_err_max_vs_fp64"] = float(np.abs(d).max())
        out[key]["phase_density_err_vs_fp64"] = [out[key]["rho_p5"] - out["f64"]["rho_p5"],
                                                 out[key]["rho_p95"] - out["f64"]["rho_p95"]]
    # Algorithmic bytes per update are 2*Q*sizeof(population) = 152 B (FP32),
    # 76 B (16-bit), psi and masks excluded; `achieved` divides them by the
    # step's kernel time from HIP events.  The row-slab one-pass sweep
    # (sc_sweep 3) moves about that; the two-pass step moves 236 B (FP32: the
    # density pass reads the 19 populations again and stores psi, the sweep
    # reads psi).  Measured HBM bytes per step for each storage's tuned form:
    # profiles/d3q19_spinodal_192_traffic.json.
    n = 192 ** 3
    ms = out["f32"]["seconds"] / steps * 1e3
    kms = out["f32"]["kernel_us_per_step"] * 1e-3
    achieved = 152.0 * n / (kms * 1e-3) / 1e9
    forms = {0: "density pass + sweep (k_density + k_step)", 1: "tiled one-pass sweep (k_step_scz)",
             2: "tiled one-pass sweep (k_step_scz)", 3: "row-slab one-pass sweep (k_step_scr)"}
    rl16 = None
    if "bf16" in out:
        k16 = out["bf16"]["kernel_us_per_step"] * 1e-3
        a16 = 76.0 * n / (k16 * 1e-3) / 1e9
        rl16 = {"storage": "bf16, stochastic rounding (the Shan-Chen default)", "bound": "hbm",
                "achieved": round(a16, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(a16 / PEAK_HBM_GBS, 4),
                "traffic": spinodal_traffic("bf16"), "bytes_per_update": 76,
                "kernel": forms.get(out["bf16"]["launch"]["sc_sweep"]), "kernel_us": out["bf16"]["kernel_us_per_step"]}
    cpu = None
    if not args.no_cpu_baseline:
        cfg, solid, init = scenes.spinodal(storage="f32")
        init = dict(init, ux=0.0, uy=0.0, uz=0.0)
        cpu = cpu_baseline(cfg, solid, init, args.cpu_steps)
    return common("D3Q19 Shan-Chen MLUPS (FP32 storage) + error of FP32/FP16/BF16 vs FP64", out["f32"]["mlups"],
                  "MLUPS", args, steps, round(ms, 5), "fp32",
                  {"workload": "D3Q19 Shan-Chen spinodal decomposition 192^3 periodic, psi=1-exp(-rho), G=-5, "
                               "tau=1, rho=0.7+U(-0.01,0.01) seed 0"}, storage=out,
                  roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": spinodal_traffic("f32"),
                            "kernel": forms.get(out["f32"]["launch"]["sc_sweep"]), "bytes_per_update": 152,
                            "kernel_us": out["f32"]["kernel_us_per_step"]},
                  roofline_16bit=rl16, cpu_baseline=cpu)


def bench_ensemble(args):
    """C4: 512 ragged D2Q9 channels in one launch per step, to residual < 1e-6."""
    import numpy as np
    from lattibol_amd import scenes, Batch
    specs = scenes.ensemble(512, seed=1234)
    b = Batch(specs)
    n = b.total_nodes
    b.step(args.warmup)
    t0 = time.perf_counter()
    b.step(args.steps)
    dt = time.perf_counter() - t0
    mlups = n * args.steps / dt / 1e6
    b.init_rest()
    t0 = time.perf_counter()
    steps, res = b.run_to_steady(tol=1e-6, k=100, max_steps=args.max_steady_steps)
    wall = time.perf_counter() - t0
    conv = res < 1e-6
    upd = float(sum(int(steps[i]) * specs[i]["nx"] * specs[i]["ny"] for i in range(len(specs))))
    ms = dt / args.steps * 1e3
    achieved = 72.0 * n / (ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads = cpu_threads()
        O.set_threads(threads)
        nodes = 0
        t0 = time.perf_counter()
        for sp in specs[:16]:
            orc = O.PullOracle(scenes.channel_config(sp), sp["solid"])
            orc.init_equilibrium(1.0, 0.0, 0.0)
            orc.run(200)
            nodes += sp["nx"] * sp["ny"] * 200
        tc = time.perf_counter() - t0
        cpu = {"value": round(nodes / tc / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port", **cpu_info(),
               "sample": "the first 16 lattices of the ensemble, 200 steps each, FP32-deviation oracle, %.1f s" % tc}
    return common("D2Q9 batched-ensemble MLUPS", round(mlups, 1), "MLUPS", args, args.steps,
                  round(dt / args.steps * 1e3, 5), "fp32",
                  {"workload": "512 D2Q9 channels, sizes {64x64,128x64,128x128,256x128}, 1-4 circles, "
                               "u_in in [0.02,0.08], tau in [0.52,0.9], one launch per step", "nodes": n},
                  roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": profile_field(
                                "d2q9_ensemble_512_traffic.json", "hbm_bytes_per_launch"),
                            "kernel": "k_batch_step (all lattices, one launch per step)", "bytes_per_update": 72},
                  cpu_baseline=cpu,
                  steady={"converged": int(conv.sum()), "of": len(specs), "wall_s": round(wall, 2),
                          "max_steps": args.max_steady_steps, "updates": upd,
                          "effective_mlups": round(upd / wall / 1e6, 1),
                          "median_steps_converged": float(np.median(steps[conv])) if conv.any() else None})


def bench_warmstart(args):
    """C5: U-Net warm start vs plain LBM from rest, time to residual < 1e-6.
    value: wall-time ratio of the warm-start pipeline with its plain fallback
    (channels the prediction does not bring to the tolerance are run again
    from rest) over the plain runs, on every channel that is steady from
    rest; evaluation.wall_ratio is the warm start alone."""
    from lattibol_amd import warmstart
    from lattibol_amd.unet import UNet
    import numpy as np
    t0 = time.perf_counter()
    data = warmstart.dataset(args.train_samples, seed=4242, max_steps=args.max_steady_steps, log=log)
    gen = time.perf_counter() - t0
    log("training set: %d/%d converged in %.1f s" % (data["converged"].sum(), args.train_samples, gen))
    net = UNet(levels=4, base=32, in_ch=3, out_ch=3, height=128, width=128, max_batch=16, seed=1)
    hist = warmstart.train(net, data, epochs=args.epochs, batch=16, lr=1e-3, log=log)
    ev = warmstart.evaluate(net, n=args.heldout, seed=777, max_steps=args.max_steady_steps, init_mode=0,
                            extra_modes=(1,), curve_eps=(0.0, 0.01, 0.03, 0.1))
    pl = ev.pop("per_lattice")
    sp, sw = np.array(pl["steps_plain"], float), np.array(pl["steps_warm"], float)
    ok = (sp > 0) & (sw > 0)
    ev["per_channel_step_ratio_median"] = float(np.median(sw[ok] / sp[ok])) if ok.any() else None
    for c in ev["curve"]:
        log("curve: eps %.2f init mode %d: step ratio %.3f, L2 median %.2e max %.2e" % (
            c["eps"], c["init_mode"], c["step_ratio"], c["l2_median"] or 0, c["l2_max"] or 0))
    return common("time-to-steady-state, warm-start / plain LBM (wall-time ratio)",
                  round(ev["with_fallback"]["wall_ratio"], 4),
                  "ratio", args, 0, None, "fp32 LBM, bf16 U-Net",
                  {"workload": "D2Q9 channels 128x128 (1-4 circles, u_in in [0.02,0.08], tau in [0.52,0.9]); "
                               "4-level 32-channel U-Net trained on the batched solver's steady fields; "
                               "%d held-out geometries" % args.heldout},
                  higher=False, train={"samples": int(data["converged"].sum()), "generation_s": round(gen, 1),
                                       "epochs": args.epochs, "best_epoch": hist["best_epoch"],
                                       "train_s": round(hist["seconds"], 1), "final_loss": hist["loss"][-1],
                                       "best_val_loss": min(hist["val_loss"]) if hist["val_loss"] else None,
                                       "mirror_augmentation": True}, evaluation=ev)


def unet_flops(levels=4, base=32, H=128, W=128, B=16, out_ch=3):
    """Dense MFMA work of one training iteration: forward, input gradients
    (every conv but the first) and weight gradients, 2 flop per MAC, the first
    conv counted with its padded 16 input channels."""
    from lattibol_amd.unet import conv_layout
    layout, _ = conv_layout(levels, base, 3, out_ch)
    convs = [c for c in layout if c[0] == "conv"]
    res = []
    for l in range(levels - 1):
        res += [l, l]
    res += [levels - 1, levels - 1]
    for l in range(levels - 2, -1, -1):
        res += [l, l]
    fwd = dgrad = 0
    for k, (kind, co, ci, _, _) in enumerate(convs):
        npix = B * (H >> res[k]) * (W >> res[k])
        f = 2.0 * npix * 9 * ci * co
        fwd += f
        if k > 0:
            dgrad += f
    head = 2.0 * B * H * W * base * out_ch
    return fwd + head, fwd + dgrad + fwd + 3 * head


def bench_unet(args):
    """U-Net training iterations (batch 16 of 128x128, synthetic inputs):
    forward + backward + Adam, all on the BF16 MFMA kernels.  The loss is
    computed on the device every iteration and read back after the timed
    loop (a per-iteration read-back adds one host synchronisation, reported
    as ms_per_step_loss_readback)."""
    import numpy as np
    from lattibol_amd.unet import UNet
    from lattibol_amd._lib import DeviceArray
    net = UNet(levels=4, base=32, in_ch=3, out_ch=3, height=128, width=128, max_batch=16, seed=1)
    rng = np.random.default_rng(0)
    xd = DeviceArray.from_numpy(rng.standard_normal((16, 128, 128, 3)).astype(np.float32))
    td = DeviceArray.from_numpy(rng.standard_normal((16, 128, 128, 3)).astype(np.float32))
    md = DeviceArray.from_numpy(np.ones((16, 128, 128), np.float32))
    yd = DeviceArray((16, 128, 128, 3), np.float32)
    for _ in range(max(1, args.warmup // 50)):
        net.forward(xd, 16); net.backward(td, md, 16); net.adam(1e-4)
    net.sync()
    steps = max(1, args.steps // 100)
    t0 = time.perf_counter()
    for _ in range(steps):
        net.forward(xd, 16); net.backward(td, md, 16, want_loss=False); net.adam(1e-4)
    net.sync()
    dt = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        net.forward(xd, 16); net.backward(td, md, 16); net.adam(1e-4)
    net.sync()
    dt_sync = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        net.forward(xd, 16, yd)
    dtf = (time.perf_counter() - t0) / steps
    ffwd, ftrain = unet_flops()
    net.close()
    return common("U-Net training iteration (batch 16 x 128x128)", round(dt * 1e3, 3), "ms", args, steps,
                  round(dt * 1e3, 3), "bf16", {"workload": "4-level 32-channel U-Net, 3x3 convs, batch 16 of 128x128, "
                                                           "synthetic inputs; forward+backward+Adam"},
                  higher=False, train_tflops=round(ftrain / dt / 1e12, 2), train_gflop=round(ftrain / 1e9, 1),
                  forward_ms=round(dtf * 1e3, 3), forward_tflops=round(ffwd / dtf / 1e12, 2),
                  ms_per_step_loss_readback=round(dt_sync * 1e3, 3),
                  mfma_peak_tflops=2500.0)


def bench_reference(args):
    """The reference's launchers (streaming_d3q19, get/copy/apply_border,
    collision_d3q19, read_value_3d) driven by the reference step loop on the
    C2 lattice: the drop-in path for callers that keep the reference's call
    sites.  Reference physics: Zou-He inlet u = 0.05 / outlet rho = 1,
    full-way bounce-back at the sphere, periodic y and z, tau = 0.56.
    Traffic of the reference's two passes per update: streaming swaps
    9 x (2 loads + 2 stores) x 4 B + 1 B mask = 145 B, collision 19 loads +
    19 stores + 4 macro planes + the plot float4 + 1 B mask = 185 B."""
    import lattibol_amd as lb
    from lattibol_amd import scenes
    out = {}
    cfg, solid, init = scenes.sphere_channel()
    nu = (1.0 / cfg["tau_inv"] - 0.5) / 3.0
    for nd in (1, 2):
        sim = lb.Simulation(solid, deviceCount=nd, ux=0.05, rho=1.0, viscosity=nu, cursor=(128, 64, 64))
        sim.runLBM(max(1, args.warmup // 4))
        sim.valueCur()
        steps = max(10, args.steps // 5)
        t0 = time.perf_counter()
        sim.runLBM(steps)
        sim.valueCur()                      # the probe read-back synchronises
        dt = time.perf_counter() - t0
        n = solid.size
        out["subdomains_%d" % nd] = {"mlups": round(n * steps / dt / 1e6, 1), "ms_per_step": round(dt / steps * 1e3, 5),
                                     "achieved_gbs_at_330B": round(330.0 * n * steps / dt / 1e9, 1)}
        sim.close()
    v = out["subdomains_1"]
    return common("D3Q19 MLUPS through the reference launchers", v["mlups"], "MLUPS", args, steps, v["ms_per_step"],
                  "fp32", {"workload": "D3Q19 256x128x128 sphere channel, reference physics and launch sequence"},
                  roofline={"bound": "hbm", "achieved": v["achieved_gbs_at_330B"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                            "frac": round(v["achieved_gbs_at_330B"] / PEAK_HBM_GBS, 4), "traffic": None,
                            "bytes_per_update": 330}, runs=out)


def bench_copy(args):
    """Plain device copy of 512 MiB: HBM bandwidth reference and the
    calibration run for the rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    import numpy as np
    from lattibol_amd import _lib
    nbytes = 512 << 20
    a = _lib.DeviceArray((nbytes,), np.uint8)
    b = _lib.DeviceArray((nbytes,), np.uint8)
    ms = C.c_float()
    reps = max(1, args.steps)
    _lib.check(_lib.lib.lb_copy_device(b.ptr, a.ptr, nbytes, reps, C.byref(ms)), "lb_copy_device")
    gbs = 2 * nbytes / (ms.value * 1e-3) / 1e9
    return common("device copy bandwidth (read + write)", round(gbs, 1), "GB/s", args, reps, round(ms.value, 5),
                  "u8", {"workload": "k_copy16, 512 MiB read + 512 MiB written per launch"},
                  bytes_read_per_launch=nbytes, bytes_written_per_launch=nbytes)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported "
              "(the z-slab halo exchange is exercised on one device)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload != "sphere":
        fn = {"cavity": bench_cavity, "spinodal": bench_spinodal, "ensemble": bench_ensemble,
              "warmstart": bench_warmstart, "unet": bench_unet, "reference": bench_reference,
              "copy": bench_copy}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return
    import numpy as np
    import lattibol_amd as lb
    from lattibol_amd import scenes

    cfg, solid, init = scenes.sphere_channel(storage=args.storage, block_x=args.block_x, xcd_remap=args.xcd_remap,
                                             mem_mode=args.mem_mode, layout=args.layout)
    nodes = cfg["nx"] * cfg["ny"] * cfg["nz"]
    esz = {"f32": 4, "f32s": 4, "f16": 2, "bf16": 2, "f64": 8}[args.storage]
    bytes_per_update = 2 * 19 * esz
    print("bench: creating solver %dx%dx%d %s" % (cfg["nx"], cfg["ny"], cfg["nz"], args.storage), file=sys.stderr,
          flush=True)
    sol = lb.Solver(cfg, solid)
    sol.init_equilibrium(init["rho"], init["ux"], init["uy"], init["uz"])
    launch = sol.launch_config
    print("bench: launch %s" % json.dumps(launch), file=sys.stderr, flush=True)
    sol.step(args.warmup)
    sol.prepare(args.steps)          # graph capture stays out of the timed region
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
    errors = sphere_errors(args.error_steps) if args.error_steps > 0 else None

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
        "dtype": {"f32": "fp32", "f32s": "fp32", "f16": "fp16-storage/fp32-compute", "bf16": "bf16-storage/fp32-compute",
                  "f64": "fp64"}[args.storage],
        "data": "synthetic",
        "config": {"workload": "D3Q19 BGK channel past a sphere 256x128x128 (r=16 at (64,64,64)), inlet u=0.05, "
                               "zero-gradient outflow, periodic y, z-slab halo planes, tau=0.56, %s" % args.storage,
                   "nodes": nodes, "bytes_per_update": bytes_per_update, "launch": launch,
                   "pct_hbm_peak_whole_step": round(100 * mlups * 1e6 * bytes_per_update / (PEAK_HBM_GBS * 1e9), 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "storage_error_vs_fp64": errors,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
