"""The benchmark configurations at their stated sizes on the GPU.

C1  D2Q9 lid-driven cavity 256x256, tau 0.6, lid u = 0.1, 10,000 steps:
    bit-identical to the oracle.
C2  D3Q19 channel past a sphere 256x128x128, 5,000 steps in FP64, FP32,
    FP16 and BF16 storage: velocity and density error against FP64.
C3  D3Q19 Shan-Chen spinodal 192^3 with FP16 / BF16 shifted storage (both
    rounding modes): 100 steps bit-identical to the oracle; 20,000 steps:
    mass drift and phase densities against the FP64 run of the same lattice.
C4  the seed-1234 ensemble of 512 D2Q9 channels run to residual < 1e-6 with
    the device-side retirement loop: counts, steady-state mass balance, and
    members bit-identical to single-lattice runs of the same step count whose
    own residual first crosses the tolerance at that step.
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

O.set_threads(min(16, os.cpu_count() or 1))


def test_c1_cavity_256_10k_steps_bit_identical(gpu):
    from lattibol_amd import scenes
    cfg, solid, init = scenes.cavity()
    print("C1 cavity 256x256, 10,000 steps", flush=True)
    sol = gpu.Solver(cfg, solid)
    orc = O.PullOracle(cfg, solid)
    sol.init_equilibrium(init["rho"], init["ux"], init["uy"])
    orc.init_equilibrium(init["rho"], init["ux"], init["uy"])
    sol.step(10000)
    orc.run(10000)
    got = sol.populations()
    assert np.count_nonzero(got != orc.A) == 0
    m, mo = sol.macros(), orc.macros()
    for k in ("rho", "ux", "uy"):
        np.testing.assert_array_equal(m[k], mo[k])
    # the lid has set the cavity in motion
    assert np.abs(m["ux"]).max() > 0.05
    sol.close()


def test_c2_sphere_storage_error_vs_fp64(gpu):
    """The channel is open (inlet density fixed, zero-gradient outflow), so
    its mass changes physically; each storage's mass is compared with the
    FP64 run's.  Measured: FP32 velocity 3.4e-5 (relative L2), density
    4.6e-6; FP16 3.5e-3 / 4.3e-4; BF16 1.4e-2 / 9.6e-4."""
    from lattibol_amd import scenes
    steps = 5000
    ref = None
    err = {}
    for st in ("f64", "f32", "f16", "bf16"):
        cfg, solid, init = scenes.sphere_channel(storage=st)
        sol = gpu.Solver(cfg, solid)
        sol.init_equilibrium(init["rho"], init["ux"], init["uy"], init["uz"])
        sol.step(steps)
        mass = sol.reduce()[0]
        m = sol.macros()
        sol.close()
        u = np.stack([m["ux"], m["uy"], m["uz"]]).astype(np.float64)
        rho = m["rho"].astype(np.float64)
        if ref is None:
            ref = (u, rho, mass)
            continue
        err[st] = (np.linalg.norm(u - ref[0]) / np.linalg.norm(ref[0]), np.abs(u - ref[0]).max(),
                   np.linalg.norm(rho - ref[1]) / np.linalg.norm(ref[1]), np.abs(rho - ref[1]).max(),
                   (mass - ref[2]) / ref[2])
        print("C2 %s vs f64 after %d steps: u L2 rel %.3e max %.3e, rho L2 rel %.3e max %.3e, mass %.2e" % (
            (st, steps) + err[st]), flush=True)
    assert err["f32"][0] < 1e-4 and err["f32"][2] < 1e-5 and abs(err["f32"][4]) < 1e-5
    assert err["f16"][0] < 1e-2 and err["f16"][2] < 2e-3
    assert err["bf16"][0] < 4e-2 and err["bf16"][2] < 4e-3


@pytest.mark.parametrize("storage,round_mode", [("f16", "default"), ("bf16", "default"), ("f16", "nearest"),
                                                ("bf16", "nearest")])
def test_c3_spinodal_192_16bit_100_steps_bit_identical(gpu, storage, round_mode):
    from lattibol_amd import scenes
    cfg, solid, init = scenes.spinodal(storage=storage, round_mode=round_mode)
    print("C3 spinodal 192^3", storage, round_mode, "100 steps", flush=True)
    sol = gpu.Solver(cfg, solid)
    orc = O.PullOracle(cfg, solid)
    sol.init_equilibrium(init["rho"], 0.0, 0.0, 0.0)
    orc.init_equilibrium(init["rho"], 0.0, 0.0, 0.0)
    sol.step(100)
    orc.run(100)
    assert np.count_nonzero(sol.populations() != orc.A) == 0
    np.testing.assert_array_equal(sol.macros()["rho"], orc.macros()["rho"])
    sol.close()


def _spinodal_run(gpu, storage, round_mode, steps):
    from lattibol_amd import scenes
    cfg, solid, init = scenes.spinodal(storage=storage, round_mode=round_mode)
    sol = gpu.Solver(cfg, solid)
    sol.init_equilibrium(init["rho"], 0.0, 0.0, 0.0)
    m0 = sol.reduce()[0]
    sol.step(steps)
    m1 = sol.reduce()[0]
    rho = sol.macros()["rho"].astype(np.float64)
    sol.close()
    p5, p95 = np.percentile(rho, [5, 95])
    return dict(drift=(m1 - m0) / m0, p5=p5, p95=p95)


def test_c3_spinodal_192_20k_steps_against_fp64(gpu):
    """Spinodal decomposition is chaotic, so after 20,000 steps each storage
    follows its own realisation of the pattern; the phase densities (5th /
    95th percentiles of rho) and the mass are what must agree with FP64."""
    steps = 20000
    print("C3 spinodal 192^3, 20,000 steps: f64, f32, f16, bf16 (stochastic), f16 / bf16 nearest", flush=True)
    ref = _spinodal_run(gpu, "f64", "default", steps)
    print("f64", ref, flush=True)
    assert abs(ref["drift"]) < 1e-12
    assert ref["p95"] - ref["p5"] > 1.5           # decomposed into two phases
    out = {}
    for st, rm in (("f32", "default"), ("f16", "default"), ("bf16", "default"), ("f16", "nearest"),
                   ("bf16", "nearest")):
        out[(st, rm)] = r = _spinodal_run(gpu, st, rm, steps)
        print(st, rm, r, flush=True)
    f32 = out[("f32", "default")]
    assert abs(f32["drift"]) < 1e-5
    assert abs(f32["p5"] - ref["p5"]) < 1e-3 and abs(f32["p95"] - ref["p95"]) < 1e-3
    # the default for 16-bit Shan-Chen storage: stochastic rounding
    for st in ("f16", "bf16"):
        r = out[(st, "default")]
        assert abs(r["drift"]) < 5e-4, (st, r)
        assert abs(r["p5"] - ref["p5"]) < 3e-3 and abs(r["p95"] - ref["p95"]) < 3e-3, (st, r)
    # why it is the default: nearest-even loses mass (FP16) or never
    # separates the phases (BF16)
    assert abs(out[("f16", "nearest")]["drift"]) > 10 * abs(out[("f16", "default")]["drift"])
    bn = out[("bf16", "nearest")]
    assert bn["p95"] - bn["p5"] < 0.1


def test_c4_ensemble_512_to_steady(gpu):
    from lattibol_amd import scenes
    specs = scenes.ensemble(512, seed=1234)
    print("C4 ensemble of 512 channels to residual < 1e-6", flush=True)
    b = gpu.Batch(specs)
    tol, k, cap = 1e-6, 100, 300000
    steps, res = b.run_to_steady(tol=tol, k=k, max_steps=cap)
    conv = res < tol
    div = np.isinf(res)
    capped = ~conv & ~div
    print("converged %d diverged %d at the cap %d; steps median %d max %d" % (
        conv.sum(), div.sum(), capped.sum(), np.median(steps[conv]), steps[conv].max()), flush=True)
    assert conv.sum() >= 400
    assert (steps % k == 0).all() and (steps[conv] < cap).all() and (steps[capped] == cap).all()
    # steady state: the mass crossing the line between columns 2 and 3
    # equals the mass crossing the line between columns nx-3 and nx-2 (both
    # clear of obstacles).  Counted from the stored post-collision populations
    # (deviations; the weights cancel): +x movers of the left column minus
    # -x movers of the right one, without the diagonals that bounce back at
    # the walls of rows 0 and ny-1 (storage rows grow downward).
    def crossing(s, x):
        r, l = s[:, :, x].astype(np.float64), s[:, :, x + 1].astype(np.float64)
        return (r[1].sum() + r[5, 1:].sum() + r[8, :-1].sum()) - (l[3].sum() + l[6, 1:].sum() + l[7, :-1].sum())
    mism = []
    for i in np.where(conv)[0]:
        s = b.populations(i)
        a, c = crossing(s, 2), crossing(s, s.shape[2] - 3)
        mism.append(abs(a - c) / abs(a))
    mism = np.array(mism)
    print("mass flux mismatch between the two lines: median %.2e max %.2e" % (np.median(mism), mism.max()),
          flush=True)
    assert mism.max() < 1e-4
    # members against single-lattice runs of the same length: bit-identical
    # populations, and the single run's residual first falls below the
    # tolerance exactly at the batch's step count
    picks = list(np.where(conv)[0][:12]) + list(np.argsort(np.where(conv, steps, 1 << 30))[:4])
    for i in sorted(set(int(p) for p in picks)):
        sp = specs[i]
        sol = gpu.Solver(scenes.channel_config(sp), sp["solid"])
        sol.init_uniform(1.0, 0.0, 0.0, 0.0)
        first = None
        n = 0
        while n < steps[i]:
            sol.step(k)
            n += k
            r = sol.residual()
            if n > k and r < tol and first is None:
                first = n
        assert first == steps[i], (i, first, steps[i])
        np.testing.assert_array_equal(b.populations(i), sol.populations()[:, 0])
        sol.close()
    b.close()
