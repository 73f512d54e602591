"""Learned warm start: the batched solver produces steady channel fields, the
U-Net learns them, and held-out channels are initialised at equilibrium with
the predicted (rho, u) and relaxed by the plain LBM to the same residual.

Residual: ||u(t) - u(t-k)||_2 / ||u(t)||_2 over fluid nodes, k = 100 steps,
FP64 reduction on the GPU (lb_batch_run_to_steady).
"""
import time

import numpy as np

from . import scenes
from ._lib import DeviceArray
from .batch import Batch
from .unet import UNet, make_inputs, make_targets


def solve_ensemble(specs, tol=1e-6, k=100, max_steps=300000, unet=None, init_mode=0, fallback_plain=True):
    """Run every channel to steady state; returns per-lattice steps, residuals,
    macros and the wall time (including an optional U-Net inference and, with
    fallback_plain, the plain runs from rest of channels the warm start did
    not bring to tol)."""
    b = Batch(specs)
    t0 = time.perf_counter()
    if unet is not None:
        x = make_inputs(np.stack([s["solid"] for s in specs]), [s["u_in"] for s in specs], [s["tau"] for s in specs])
        n = len(specs)
        H, W = x.shape[1:3]
        bsz = unet.desc.max_batch
        xd = DeviceArray((bsz, H, W, 3), np.float32)
        yd = DeviceArray((bsz, H, W, 3), np.float32)
        ud = DeviceArray((bsz,), np.float32)
        fields = DeviceArray((n, 3, H, W), np.float32)
        for s in range(0, n, bsz):
            m = min(bsz, n - s)
            xb = np.zeros((bsz, H, W, 3), np.float32)
            xb[:m] = x[s:s + m]
            xd.upload(xb)
            uin = np.zeros(bsz, np.float32)
            uin[:m] = [sp["u_in"] for sp in specs[s:s + m]]
            ud.upload(uin)
            unet.forward(xd, m, yd)
            # predicted planes go straight into the packed field array (device)
            sub = _Offset(fields, s * 3 * H * W * 4)
            unet.to_fields(yd, m, ud, sub)
        b.init_fields_device(fields.ptr.value, init_mode)
    t_init = time.perf_counter() - t0
    steps, res = b.run_to_steady(tol=tol, k=k, max_steps=max_steps)
    macros = [b.macros(i) for i in range(len(specs))]
    b.close()
    fallback = []
    if unet is not None and fallback_plain:
        # the plain LBM is the fallback: a channel that did not reach tol from
        # the predicted field (diverged, or still moving at max_steps) is run
        # again from rest; its steps and wall time count for the warm start
        fallback = [int(i) for i in np.where(~(res < tol))[0]]
        if fallback:
            again = solve_ensemble([specs[i] for i in fallback], tol=tol, k=k, max_steps=max_steps)
            steps = np.array(steps, copy=True)
            res = np.array(res, copy=True)
            for j, i in enumerate(fallback):
                steps[i] += again["steps"][j]
                res[i] = again["res"][j]
                macros[i] = again["macros"][j]
    wall = time.perf_counter() - t0
    return dict(steps=steps, res=res, macros=macros, wall=wall, t_init=t_init, fallback=fallback)


class _Offset:
    """A view of a DeviceArray at a byte offset (for to_fields)."""

    def __init__(self, arr, byte_off):
        import ctypes as C
        self.ptr = C.c_void_p(arr.ptr.value + byte_off)


def dataset(n, seed, size=(128, 128), tol=1e-6, k=100, max_steps=300000, chunk=2048, log=None):
    """n seeded channels run to steady state in batches of `chunk` lattices."""
    specs = scenes.ensemble(n, seed=seed, size=size)
    parts = []
    for s in range(0, n, chunk):
        parts.append(solve_ensemble(specs[s:s + chunk], tol=tol, k=k, max_steps=max_steps))
        if log:
            log("dataset: %d / %d channels solved" % (min(n, s + chunk), n))
    out = dict(res=np.concatenate([p["res"] for p in parts]), steps=np.concatenate([p["steps"] for p in parts]),
               macros=[m for p in parts for m in p["macros"]], wall=sum(p["wall"] for p in parts))
    ok = out["res"] < tol
    solid = np.stack([s["solid"] for s in specs])
    u_in = np.array([s["u_in"] for s in specs])
    tau = np.array([s["tau"] for s in specs])
    rho = np.stack([m["rho"] for m in out["macros"]])
    ux = np.stack([m["ux"] for m in out["macros"]])
    uy = np.stack([m["uy"] for m in out["macros"]])
    t = make_targets(rho, ux, uy, u_in)
    fluid = (1 - solid).astype(bool)
    ok = ok & np.array([np.isfinite(t[i][fluid[i]]).all() for i in range(len(specs))])
    return dict(specs=specs, converged=ok, steps=out["steps"], wall=out["wall"],
                x=make_inputs(solid, u_in, tau), t=np.where(np.isfinite(t), t, 0).astype(np.float32),
                mask=(1 - solid).astype(np.float32), u_in=u_in, tau=tau)


def fit_stats(t, mask):
    m = mask.astype(bool)
    st = []
    for c in range(t.shape[-1]):
        v = t[..., c][m]
        st += [float(v.mean()), float(v.std() + 1e-8)]
    return np.array(st, np.float32)


def standardise(t, stats):
    out = np.empty_like(t)
    for c in range(t.shape[-1]):
        out[..., c] = (t[..., c] - stats[2 * c]) / stats[2 * c + 1]
    return out


def _mirror(x, t, m):
    """The channel reflected about its centre line (rows reversed): the same
    boundary value problem with uy -> -uy (inputs, targets, masks)."""
    x = x[:, ::-1]
    t = t[:, ::-1].copy()
    t[..., 1] = -t[..., 1]
    return x, t, m[:, ::-1]


def train(net, data, epochs=60, batch=16, lr=1e-3, seed=0, log=None, mirror=True, val_frac=0.05):
    """Adam with a cosine learning rate on the converged samples.  mirror:
    every batch reflects a random half of its channels about the centre line
    (the problem's symmetry).  A held-out fraction val_frac tracks the
    validation loss after every epoch; the parameters of the best epoch are
    kept (early stopping)."""
    ok = data["converged"]
    x, t, m = data["x"][ok], data["t"][ok], data["mask"][ok]
    net.stats = fit_stats(t, m)
    rng = np.random.default_rng(seed)
    n_all = x.shape[0]
    perm = rng.permutation(n_all)
    # a validation set of at least one batch, only for sets of 10+ batches
    nval = max(batch, int(round(val_frac * n_all))) if n_all >= 10 * batch and val_frac > 0 else 0
    vidx, tidx = perm[:nval], perm[nval:]
    n = tidx.size
    H, W = x.shape[1:3]
    xd = DeviceArray((batch, H, W, 3), np.float32)
    td = DeviceArray((batch, H, W, 3), np.float32)
    md = DeviceArray((batch, H, W), np.float32)
    yd = DeviceArray((batch, H, W, 3), np.float32)
    hist, vhist = [], []
    best = (np.inf, None, -1)
    t0 = time.perf_counter()
    for ep in range(epochs):
        order = tidx[rng.permutation(n)]
        tot = 0.0
        nb = 0
        lr_ep = lr * 0.5 * (1 + np.cos(np.pi * ep / epochs))
        for s in range(0, n - batch + 1, batch):
            idx = order[s:s + batch]
            xb, tb, mb = x[idx], t[idx], m[idx]
            if mirror:
                flip = rng.random(batch) < 0.5
                if flip.any():
                    xf, tf, mf = _mirror(xb[flip], tb[flip], mb[flip])
                    xb, tb, mb = xb.copy(), tb.copy(), mb.copy()
                    xb[flip], tb[flip], mb[flip] = xf, tf, mf
            xd.upload(xb)
            td.upload(standardise(tb, net.stats))
            md.upload(mb)
            net.forward(xd, batch)
            tot += net.backward(td, md, batch)
            net.adam(lr=float(lr_ep))
            nb += 1
        hist.append(tot / max(nb, 1))
        if nval:
            vl, vb = 0.0, 0
            for s in range(0, nval - batch + 1, batch):
                idx = vidx[s:s + batch]
                xd.upload(x[idx])
                net.forward(xd, batch, yd)
                d = yd.numpy() - standardise(t[idx], net.stats)
                w = m[idx][..., None]
                vl += float((d * d * w).sum() / max(w.sum() * 3, 1.0))
                vb += 1
            vhist.append(vl / max(vb, 1))
            if vhist[-1] < best[0]:
                best = (vhist[-1], net.params(), ep)
        if log:
            log("epoch %d loss %.5f%s (%.1f s)" % (ep, hist[-1], (" val %.5f" % vhist[-1]) if vhist else "",
                                                    time.perf_counter() - t0))
    if best[1] is not None:
        net.set_params(best[1])
    net.sync()
    return dict(loss=hist, val_loss=vhist, best_epoch=best[2], seconds=time.perf_counter() - t0, samples=n)


def velocity_l2(a, b, mask):
    """||u_a - u_b|| / ||u_b|| over fluid nodes."""
    m = mask.astype(bool)
    da = np.concatenate([(a["ux"] - b["ux"])[m], (a["uy"] - b["uy"])[m]]).astype(np.float64)
    nb = np.concatenate([b["ux"][m], b["uy"][m]]).astype(np.float64)
    return float(np.linalg.norm(da) / max(np.linalg.norm(nb), 1e-30))


def smooth_mode(shape, seed):
    """A smooth O(1) pattern (product of two half sines with seeded
    wavenumbers and phases) for controlled perturbations of a field."""
    H, W = shape
    y, x = np.mgrid[0:H, 0:W]
    a = np.random.default_rng(seed).uniform(0.5, 1.5, 3)
    return np.sin(np.pi * (a[0] * x / W + 0.3)) * np.sin(np.pi * (a[1] * y / H + 0.2)) * a[2]


def relax_from_fields(specs, fields, tol=1e-6, k=100, max_steps=300000, init_mode=0):
    """Initialise every channel from host fields (dicts of rho, ux, uy) and
    relax to tol; returns steps, residuals, macros and wall time."""
    b = Batch(specs)
    arr = np.stack([np.stack([f["rho"], f["ux"], f["uy"]]) for f in fields]).astype(np.float32)
    d = DeviceArray(arr.shape, np.float32)
    d.upload(arr)
    t0 = time.perf_counter()
    b.init_fields_device(d.ptr.value, init_mode)
    steps, res = b.run_to_steady(tol=tol, k=k, max_steps=max_steps)
    wall = time.perf_counter() - t0
    macros = [b.macros(i) for i in range(len(specs))]
    b.close()
    return dict(steps=steps, res=res, macros=macros, wall=wall)


def error_curve(specs, plain, eps_list=(0.0, 0.01, 0.03, 0.1), modes=(0, 1), tol=1e-6, k=100, max_steps=300000):
    """What a warm start of a given accuracy buys: every channel starts from
    its own plain steady field (rho - 1 and u) scaled by (1 + eps * a smooth
    mode) and relaxes to tol; step ratio over the plain run from rest and
    velocity L2 against the plain steady field, per eps and init mode.  eps
    = 0 is the floor of any equilibrium (mode 0) or Chapman-Enskog (mode 1)
    warm start."""
    out = []
    for mode in modes:
        for eps in eps_list:
            fields = []
            for i, f in enumerate(plain["macros"]):
                sm = smooth_mode(f["rho"].shape, i)
                fields.append(dict(rho=1 + (f["rho"] - 1) * (1 + eps * sm), ux=f["ux"] * (1 + eps * sm),
                                   uy=f["uy"] * (1 + eps * sm)))
            r = relax_from_fields(specs, fields, tol=tol, k=k, max_steps=max_steps, init_mode=mode)
            both = (r["res"] < tol) & (plain["res"] < tol)
            l2 = np.array([velocity_l2(r["macros"][i], plain["macros"][i], 1 - specs[i]["solid"])
                           for i in range(len(specs))])
            out.append(dict(eps=eps, init_mode=mode, converged=int(both.sum()),
                            step_ratio=float(r["steps"][both].sum() / max(plain["steps"][both].sum(), 1)),
                            l2_median=float(np.median(l2[both])) if both.any() else None,
                            l2_max=float(l2[both].max()) if both.any() else None))
    return out


def evaluate(net, n=256, seed=777, tol=1e-6, k=100, max_steps=300000, size=(128, 128), init_mode=0, extra_modes=(),
             curve_eps=()):
    """Warm start vs plain start from rest on n held-out channels.

    Time to steady state exists only for channels whose flow has a steady
    state: a first plain run finds them (residual < tol within max_steps; the
    rest shed vortices or diverge), then the plain start from rest and the
    warm start (U-Net inference + initialisation + relaxation) are timed on
    exactly that set."""
    specs_all = scenes.ensemble(n, seed=seed, size=size)
    probe = solve_ensemble(specs_all, tol=tol, k=k, max_steps=max_steps)
    steady = np.where(probe["res"] < tol)[0]
    specs = [specs_all[i] for i in steady]
    m = len(specs)
    plain = solve_ensemble(specs, tol=tol, k=k, max_steps=max_steps)
    warm = solve_ensemble(specs, tol=tol, k=k, max_steps=max_steps, unet=net, init_mode=init_mode,
                          fallback_plain=False)
    warm_fb = solve_ensemble(specs, tol=tol, k=k, max_steps=max_steps, unet=net, init_mode=init_mode)
    both = (plain["res"] < tol) & (warm["res"] < tol)
    l2 = np.array([velocity_l2(warm["macros"][i], plain["macros"][i], 1 - specs[i]["solid"]) for i in range(m)])
    # prediction error of the network itself (before relaxation)
    x = make_inputs(np.stack([s["solid"] for s in specs]), [s["u_in"] for s in specs], [s["tau"] for s in specs])
    y = net.predict(x)
    pred_err = []
    for i in range(m):
        u = specs[i]["u_in"]
        pr = dict(ux=u * (y[i, ..., 0] * net.stats[1] + net.stats[0]), uy=u * (y[i, ..., 1] * net.stats[3] + net.stats[2]))
        pred_err.append(velocity_l2(pr, plain["macros"][i], 1 - specs[i]["solid"]))
    pred_err = np.array(pred_err)
    curve = error_curve(specs, plain, eps_list=curve_eps, tol=tol, k=k, max_steps=max_steps) if curve_eps else []
    extra = {}
    for mode in extra_modes:
        w2 = solve_ensemble(specs, tol=tol, k=k, max_steps=max_steps, unet=net, init_mode=mode, fallback_plain=False)
        b2 = (plain["res"] < tol) & (w2["res"] < tol)
        l2b = np.array([velocity_l2(w2["macros"][i], plain["macros"][i], 1 - specs[i]["solid"]) for i in range(m)])
        extra["init_mode_%d" % mode] = dict(
            converged_warm=int((w2["res"] < tol).sum()), converged_both=int(b2.sum()),
            step_ratio=float(w2["steps"][b2].sum() / max(plain["steps"][b2].sum(), 1)),
            wall_warm=w2["wall"], wall_ratio=float(w2["wall"] / plain["wall"]),
            l2_max=float(l2b[b2].max()) if b2.any() else None)
    return dict(extra=extra, curve=curve,
        n=n, steady=m, unsteady_or_diverged=n - m, init_mode=init_mode,
        converged_plain=int((plain["res"] < tol).sum()), converged_warm=int((warm["res"] < tol).sum()),
        converged_both=int(both.sum()),
        steps_plain=int(plain["steps"][both].sum()), steps_warm=int(warm["steps"][both].sum()),
        step_ratio=float(warm["steps"][both].sum() / max(plain["steps"][both].sum(), 1)),
        wall_plain=plain["wall"], wall_warm=warm["wall"], wall_ratio=float(warm["wall"] / plain["wall"]),
        inference_s=warm["t_init"],
        # channels steady from rest that did not reach tol from the warm start:
        # residual +inf = retired by the divergence guard, finite = step cap
        warm_not_converged=[dict(index=int(i), steps=int(warm["steps"][i]), residual=float(warm["res"][i]))
                            for i in np.where((plain["res"] < tol) & ~(warm["res"] < tol))[0]],
        # the warm start with the plain fallback, over every steady channel
        with_fallback=dict(fallback=len(warm_fb["fallback"]), converged=int((warm_fb["res"] < tol).sum()),
                           step_ratio=float(warm_fb["steps"][plain["res"] < tol].sum() /
                                            max(plain["steps"][plain["res"] < tol].sum(), 1)),
                           wall_warm=warm_fb["wall"], wall_ratio=float(warm_fb["wall"] / plain["wall"])),
        l2_max=float(l2[both].max()) if both.any() else None, l2_median=float(np.median(l2[both])) if both.any() else None,
        pred_err_median=float(np.median(pred_err[both])) if both.any() else None,
        per_lattice=dict(steps_plain=plain["steps"].tolist(), steps_warm=warm["steps"].tolist(),
                         l2=l2.tolist(), pred_err=pred_err.tolist()),
    )
