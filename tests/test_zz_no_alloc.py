"""lb_conv3x3_bf16_noalloc (the convolution entry point that allocates
nothing: its K-split shares live in a fixed per-GPU device array) against
lb_conv3x3_bf16 (which allocates them): bit-identical outputs on two calls
in a row, the second smaller, in both the BF16 bias+ReLU and the FP32 forms;
a call whose shares exceed the array is rejected before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH_FLOATS = 4 << 20


def _split_shares(B, H, W, ci, co):
    """K-split share count of the 16-pixel tile plan (the library's rule)."""
    nb = 2 if co % 64 == 0 else 1
    if W % 32 == 0 and H % 4 == 0:
        return 1
    if not (W % 16 == 0 and H % 8 == 0):
        return 1
    wgs = B * (H // 8) * (W // 16) * (co // (32 * nb))
    ks = 1
    while ks * 2 <= ci // 32 and wgs * ks < 512:
        ks *= 2
    return ks


def _run(lib, fn, mode, B, H, W, ci, co, seed):
    from lattibol_amd import _lib
    rng = np.random.default_rng(seed)
    bf = lambda a: (np.asarray(a, np.float32).view(np.uint32) >> 16).astype(np.uint16)   # truncation: any BF16
    w = _lib.DeviceArray.from_numpy(bf(rng.standard_normal((co, 9, ci)) * 0.1))
    if mode == 0:
        x = _lib.DeviceArray.from_numpy(bf(rng.standard_normal((B, H, W, ci))))
        b = _lib.DeviceArray.from_numpy(rng.standard_normal(co).astype(np.float32))
        y = _lib.DeviceArray((B, H, W, co), np.uint16)
        bptr = b.ptr
    else:
        x = _lib.DeviceArray.from_numpy(rng.standard_normal((B, H, W, ci)).astype(np.float32))
        y = _lib.DeviceArray((B, H, W, co), np.float32)
        bptr = None
    rc = getattr(lib, fn)(mode, x.ptr, w.ptr, bptr, y.ptr, B, H, W, ci, co, None)
    _lib.check(rc, fn)
    _lib.check(lib.lb_device_synchronize(), "sync")
    return y.numpy()


@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_conv_equals_allocating_conv_twice(gpu, mode):
    from lattibol_amd import _lib
    lib = _lib.lib
    # the U-Net's 16x16 level at batch 16 (the largest split that fits), then
    # a smaller call reusing the same array
    for B, H, W, ci, co, seed in ((16, 16, 16, 256, 256, 1), (3, 16, 16, 128, 64, 2)):
        ks = _split_shares(B, H, W, ci, co)
        assert ks > 1 and ks * B * H * W * co <= SCRATCH_FLOATS
        print("noalloc conv mode", mode, (B, H, W, ci, co), "shares", ks, flush=True)
        got = _run(lib, "lb_conv3x3_bf16_noalloc", mode, B, H, W, ci, co, seed)
        want = _run(lib, "lb_conv3x3_bf16", mode, B, H, W, ci, co, seed)
        np.testing.assert_array_equal(got, want)
        assert np.abs(got.astype(np.float64)).max() > 0


def test_noalloc_conv_rejects_oversize_and_runs_unsplit_shapes(gpu):
    from lattibol_amd import _lib
    lib = _lib.lib
    B, H, W, ci, co = 36, 16, 16, 256, 256          # 2 shares x 36 x 256 x 256 floats
    assert _split_shares(B, H, W, ci, co) * B * H * W * co > SCRATCH_FLOATS
    rc = lib.lb_conv3x3_bf16_noalloc(0, 1, 1, 1, 1, B, H, W, ci, co, None)   # rejected before any launch
    assert rc == _lib.LB_ERR_SCRATCH, rc
    # a 32-pixel-segment shape needs no shares at all
    got = _run(lib, "lb_conv3x3_bf16_noalloc", 0, 2, 32, 32, 64, 64, 5)
    want = _run(lib, "lb_conv3x3_bf16", 0, 2, 32, 32, 64, 64, 5)
    np.testing.assert_array_equal(got, want)
