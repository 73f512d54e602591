"""Writes the golden fixtures of tests/golden/*.npz from the CPU oracle.

The reference ships no expected outputs (no golden vectors, KATs or
fixtures), so these small cases are produced by the oracle, committed, and
then pin two things: the oracle itself against later edits (CPU test) and
the HIP path against the same bytes (GPU test).  Each file holds the inputs
(configuration as JSON text, solid mask, initial fields) and the outputs
(raw populations after the run, macros; for the reference driver also rho,
ux, the plot buffer and the probe value).

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "lattibol_0.3.31f_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def deck_solids(shape):
    s = np.zeros(shape, np.uint8)
    if len(shape) == 2:
        H, W = shape
        s[0, :] = s[-1, :] = 1
        s[3:5, 8:11] = 1
    else:
        D, H, W = shape
        s[:, 0, :] = s[:, -1, :] = 1
        s[1:3, 3:5, 6:9] = 1
    return s


# Reference-driver decks: (shape, deviceCount, kwargs, steps, cursor)
REF_CASES = {
    "ref_d2q9_zouhe_32x8": ((8, 32), 1, dict(ux=0.05, rho=1.0, viscosity=0.05), 40, (3, 5, 0)),
    "ref_d2q9_2dev_32x8": ((8, 32), 2, dict(ux=0.04, rho=1.0, viscosity=0.08), 30, (5, 4, 0)),
    "ref_d2q9_shanchen_32x8": ((8, 32), 1, dict(ux=0.0, rho=200.0, viscosity=1 / 6, G=-120.0, Gads=-100.0), 30,
                               (4, 3, 0)),
    "ref_d2q9_force_32x8": ((8, 32), 1, dict(ux=0.0, rho=1.0, viscosity=0.1, Fx=2.0e-5), 40, (6, 4, 0)),
    "ref_d3q19_2dev_32x8x4": ((4, 8, 32), 2, dict(ux=0.03, rho=1.0, viscosity=0.05), 16, (3, 4, 1)),
}


def pull_cases():
    from lattibol_amd import scenes
    out = {}
    for st, steps in (("f32", 200), ("f16", 100), ("bf16", 100), ("f64", 100), ("f32s", 100)):
        cfg, solid, init = scenes.cavity(n=16, storage=st)
        out["pull_cavity16_%s" % st] = (cfg, solid, init, steps)
    cfg, solid, init = scenes.sphere_channel(16, 8, 6, radius=2, center=(5, 4, 3))
    out["pull_sphere16x8x6_halo_f32"] = (cfg, solid, init, 60)
    cfg, solid, init = scenes.sphere_channel(16, 8, 6, radius=2, center=(5, 4, 3), storage="bf16")
    out["pull_sphere16x8x6_halo_bf16"] = (cfg, solid, init, 60)
    cfg, solid, init = scenes.spinodal(n=10, storage="f32")
    out["pull_spinodal10_f32"] = (cfg, solid, init, 40)
    cfg, solid, init = scenes.spinodal(n=10, storage="f16", round_mode="nearest")
    out["pull_spinodal10_f16"] = (cfg, solid, init, 40)
    cfg, solid, init = scenes.spinodal(n=10, storage="f16")          # default: stochastic rounding
    out["pull_spinodal10_f16_sr"] = (cfg, solid, init, 40)
    spec = scenes.channel_spec(3, 1, (48, 24))
    cfg = scenes.channel_config(spec)
    out["pull_channel48x24_zouhe_f32s"] = (cfg, spec["solid"][None], dict(rho=1.0, ux=0.0, uy=0.0, uz=0.0), 200)
    return out


def make_ref(name, case):
    import oracle as O
    shape, nd, kw, steps, cursor = case
    s = deck_solids(shape)
    ref = O.RefOracle(s, deviceCount=nd, cursor=cursor, **kw)
    ref.runLBM(steps)
    meta = dict(kind="ref", shape=list(shape), deviceCount=nd, kwargs=kw, steps=steps, cursor=list(cursor))
    return dict(meta=np.array(json.dumps(meta)), solid=s, f=ref.f, rho=ref.rho, ux=ref.ux, plot=ref.plot,
                value_cur=ref.value_cur)


def make_pull(name, case):
    import oracle as O
    cfg, solid, init, steps = case
    orc = O.PullOracle(cfg, solid)
    shape = orc.shape
    fields = np.stack([np.broadcast_to(np.asarray(init.get(k, 0.0), np.float32), shape)
                       for k in ("rho", "ux", "uy", "uz")])
    orc.init_equilibrium(*fields)
    orc.run(steps)
    meta = dict(kind="pull", cfg=cfg, steps=steps)
    return dict(meta=np.array(json.dumps(meta)), solid=np.ascontiguousarray(orc.solid), fields=fields, pops=orc.A,
                macros=orc.macros_arr)


def main():
    import oracle as O
    O.set_threads(1)
    for name, case in REF_CASES.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **make_ref(name, case))
        print("wrote", name)
    for name, case in pull_cases().items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **make_pull(name, case))
        print("wrote", name)


if __name__ == "__main__":
    main()
