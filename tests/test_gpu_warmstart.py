"""End-to-end learned warm start on a small problem: the batched solver makes
steady fields, the U-Net trains on them on the GPU, a held-out channel set is
started from the prediction and relaxed to the same residual; the relaxed
field must match the plain run's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_small_warmstart_pipeline(gpu):
    from lattibol_amd import warmstart
    from lattibol_amd.unet import UNet
    size = (64, 64)
    data = warmstart.dataset(48, seed=31, size=size, tol=1e-5, max_steps=60000)
    ok = data["converged"]
    print("converged training channels:", int(ok.sum()), "of", ok.size, flush=True)
    assert ok.sum() >= 16
    net = UNet(levels=4, base=32, in_ch=3, out_ch=3, height=64, width=64, max_batch=8, seed=3)
    hist = warmstart.train(net, data, epochs=25, batch=8, lr=2e-3)
    assert hist["loss"][-1] < 0.5 * hist["loss"][0], hist["loss"]
    # evaluation at the production tolerance (residual 1e-6 per 100 steps)
    ev = warmstart.evaluate(net, n=8, seed=5, tol=1e-6, max_steps=200000, size=size)
    print({k: v for k, v in ev.items() if k != "per_lattice"}, ev["per_lattice"]["l2"], flush=True)
    assert ev["converged_both"] >= 4
    # with the plain fallback every channel that is steady from rest converges
    assert ev["with_fallback"]["converged"] == ev["converged_plain"]
    # The relaxed warm-started field is the plain steady field.  The stopping
    # rule bounds the change per 100 steps, not the distance to the steady
    # state: the slowest channels stop up to ~2e-2 (velocity L2) apart
    # (measured), typical ones within a few 1e-3.
    assert ev["l2_median"] < 1e-2, ev["l2_median"]
    assert ev["l2_max"] < 5e-2, ev["l2_max"]
    net.close()
