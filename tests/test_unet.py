"""U-Net surrogate: the CPU oracle pinned by finite differences and torch's own
conv2d; the GPU kernels (conv on BF16 matrix cores, full forward/backward,
Adam) checked against the oracle."""
import numpy as np
import pytest
import torch

import unet_oracle as UO
from lattibol_amd.unet import conv_layout

SMALL = dict(levels=3, base=32, in_ch=3, out_ch=3)


def _rand_params(levels, base, in_ch, out_ch, seed=0, scale=0.15):
    layout, n = conv_layout(levels, base, in_ch, out_ch)
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(n) * scale
    # the padded input channels of the first conv carry zero weights
    kind, co, ci, w, b = layout[0]
    W = p[w:b].reshape(co, 9, ci)
    W[:, :, in_ch:] = 0
    p[w:b] = W.ravel()
    return p.astype(np.float32), layout


def _he_params(levels, base, in_ch, out_ch, seed=0):
    """He-scaled random parameters (std sqrt(2 / fan_in)), so activations keep
    their scale through all levels of the production network."""
    layout, n = conv_layout(levels, base, in_ch, out_ch)
    rng = np.random.default_rng(seed)
    p = np.zeros(n)
    for i, (kind, co, ci, w, b) in enumerate(layout):
        fan = (9 * (in_ch if i == 0 else ci)) if kind == "conv" else ci
        p[w:b] = rng.standard_normal(b - w) * np.sqrt(2.0 / fan)
        p[b:b + co] = rng.standard_normal(co) * 0.05
    kind, co, ci, w, b = layout[0]
    W = p[w:b].reshape(co, 9, ci)
    W[:, :, in_ch:] = 0
    return p.astype(np.float32), layout


def test_oracle_layout_matches_binding():
    convs, _ = UO.conv_layout(4, 32, 3, 3)
    layout, n = conv_layout(4, 32, 3, 3)
    assert [(ci, co) for ci, co in convs] == [(l[2], l[1]) for l in layout[:-1]]
    # the 4-level, 32-channel network of the warm start: ~5.3 GFLOP per 128x128 sample
    flops = 0
    H = 128
    for (ci, co), l in zip(convs, [0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0]):
        flops += 2 * (H >> l) ** 2 * 9 * ci * co
    assert 5.0e9 < flops < 5.8e9


def test_oracle_gradient_matches_finite_differences():
    torch.manual_seed(0)
    p, _ = _rand_params(**SMALL, seed=1)
    params = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    x = torch.rand(1, 8, 8, 3, dtype=torch.float64)
    t = torch.rand(1, 8, 8, 3, dtype=torch.float64)
    m = (torch.rand(1, 8, 8) > 0.2).double()
    loss = UO.masked_mse(UO.forward(params, x, **SMALL, emulate=False), t, m)
    (g,) = torch.autograd.grad(loss, params)
    rng = np.random.default_rng(2)
    idx = rng.choice(p.size, 40, replace=False)
    h = 1e-6
    for i in idx:
        e = torch.zeros_like(params)
        e[i] = h
        with torch.no_grad():
            lp = UO.masked_mse(UO.forward(params + e, x, **SMALL, emulate=False), t, m)
            lm = UO.masked_mse(UO.forward(params - e, x, **SMALL, emulate=False), t, m)
        fd = (lp - lm) / (2 * h)
        assert abs(fd.item() - g[i].item()) <= 1e-6 + 1e-5 * abs(fd.item()), (i, fd.item(), g[i].item())


def test_oracle_gradient_rounding_emulation():
    # the rounding node: identity forward, BF16 nearest-even gradient
    t = torch.linspace(-3, 3, 257, dtype=torch.float64, requires_grad=True)
    c = torch.randn(257, dtype=torch.float64) * 1.2345
    (g,) = torch.autograd.grad((c * UO._GradBF16.apply(t)).sum(), t)
    assert torch.equal(g, c.to(torch.bfloat16).to(torch.float64))
    # in the network: the last conv's bias gradient sums the unrounded
    # pre-activation gradient (identical with and without the emulation);
    # its weight gradient is the product with the rounded one (within BF16's
    # half-ulp); earlier layers inherit the rounding
    p, layout = _rand_params(**SMALL, seed=3)
    x = torch.rand(2, 8, 8, 3, dtype=torch.float64)
    tg = torch.rand(2, 8, 8, 3, dtype=torch.float64)
    m = torch.ones(2, 8, 8, dtype=torch.float64)
    grads = []
    for eg in (False, True):
        params = torch.tensor(p, dtype=torch.float64, requires_grad=True)
        loss = UO.masked_mse(UO.forward(params, x, **SMALL, emulate_grad=eg), tg, m)
        grads.append(torch.autograd.grad(loss, params)[0].numpy())
    kind, co, ci, w, b = layout[-2]
    assert np.array_equal(grads[0][b:b + co], grads[1][b:b + co])
    rel = np.linalg.norm(grads[0][w:b] - grads[1][w:b]) / np.linalg.norm(grads[0][w:b])
    assert 0 < rel < 2.0 ** -8, rel
    kind, co, ci, w, b = layout[0]
    assert not np.array_equal(grads[0][b:b + co], grads[1][b:b + co])


def test_conv_reference_matches_direct_sum():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 5, 6, 16))
    w = rng.standard_normal((32, 9, 16))
    y = UO.conv3x3_ref(x, w)
    n, h, ww, co = 0, 2, 0, 7
    s = 0.0
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        hh, wx = h + ky - 1, ww + kx - 1
        if 0 <= hh < 5 and 0 <= wx < 6:
            s += float(np.dot(x[n, hh, wx], w[co, tap]))
    assert abs(y[n, h, ww, co] - s) < 1e-12


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _bf16_round(a):
    return torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


# (ci, co, B, H, W): ci % 32 == 0 takes the LDS-tile kernel in modes 0 / 1,
# compared with the gathering kernel (modes 2 / 3) too: W % 32 == 0 in all
# four tile shapes (4 or 8 rows x 32 or 64 channels); W % 16 == 0 with
# 16-pixel segments, K split over 2, 8 (uneven shares of 3 / 12 stages) or
# no workgroups; ci = 16 runs the gathering kernel only.
CONV_CASES = [(16, 32, 2, 16, 16), (32, 64, 2, 16, 16), (96, 32, 2, 16, 16), (384, 128, 2, 16, 16),
              (32, 64, 2, 8, 64), (96, 32, 1, 16, 32), (384, 128, 1, 4, 32), (64, 64, 3, 12, 32),
              (192, 64, 16, 32, 32), (32, 32, 4, 128, 128), (256, 256, 16, 16, 16), (128, 256, 16, 16, 16),
              (256, 128, 16, 16, 16), (64, 96, 2, 8, 48)]


@pytest.mark.gpu
@pytest.mark.parametrize("ci,co,B,H,W", CONV_CASES)
def test_gpu_conv3x3_matches_reference(gpu, ci, co, B, H, W):
    from lattibol_amd import _lib
    rng = np.random.default_rng(ci * 1000 + co + H * W)
    x = _bf16_round(rng.standard_normal((B, H, W, ci)))
    w = _bf16_round(rng.standard_normal((co, 9, ci)) * 0.1)
    bias = rng.standard_normal(co).astype(np.float32) * 0.1
    to_u16 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).view(torch.int16).numpy()
    xd = _lib.DeviceArray.from_numpy(to_u16(x))
    wd = _lib.DeviceArray.from_numpy(to_u16(w))
    bd = _lib.DeviceArray.from_numpy(bias)
    ref = UO.conv3x3_ref(x, w, bias, relu=True)
    xf = rng.standard_normal((B, H, W, ci)).astype(np.float32)
    xfd = _lib.DeviceArray.from_numpy(xf)
    reff = UO.conv3x3_ref(_bf16_round(xf), w)
    outs = {}
    for mode in (0, 2):
        yd = _lib.DeviceArray((B, H, W, co), np.int16)
        _lib.check(_lib.lib.lb_conv3x3_bf16(mode, xd.ptr, wd.ptr, bd.ptr, yd.ptr, B, H, W, ci, co, None), "conv")
        _lib.check(_lib.lib.lb_device_synchronize(), "sync")
        y = torch.from_numpy(yd.numpy()).view(torch.bfloat16).to(torch.float32).numpy()
        # one BF16 rounding of an FP32-accumulated value: <= 2^-8 relative
        err = np.abs(y - ref)
        assert (err <= 2.0 ** -8 * np.abs(ref) + 1e-6).all(), (mode, err.max())
        outs[mode] = y
    # FP32 in / FP32 out (the backward-data form)
    for mode in (1, 3):
        yfd = _lib.DeviceArray((B, H, W, co), np.float32)
        _lib.check(_lib.lib.lb_conv3x3_bf16(mode, xfd.ptr, wd.ptr, None, yfd.ptr, B, H, W, ci, co, None), "conv")
        _lib.check(_lib.lib.lb_device_synchronize(), "sync")
        np.testing.assert_allclose(yfd.numpy(), reff, rtol=1e-5, atol=1e-5, err_msg="mode %d" % mode)
        outs[mode] = yfd.numpy()
    # the two kernels differ only in FP32 summation order
    np.testing.assert_allclose(outs[1], outs[3], rtol=1e-5, atol=1e-5)
    assert np.abs(outs[0] - outs[2]).max() <= 2.0 ** -7 * np.abs(ref).max()


@pytest.mark.gpu
def test_gpu_bf16_conversion_instruction_matches_bit_rounding(gpu):
    # every non-NaN float: v_cvt_pk_bf16_f32 == round-to-nearest-even on the
    # bit pattern (the oracle's formula), so BF16 stores stay bit-exact
    import ctypes as C
    from lattibol_amd import _lib
    bad = C.c_uint64(123)
    _lib.check(_lib.lib.lb_check_bf16_rounding(C.byref(bad)), "lb_check_bf16_rounding")
    assert bad.value == 0


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 0xC0FFEE])
def test_gpu_stochastic_rounding_instructions_match_bit_formulas(gpu, seed):
    # every FP32 pattern with two random words each: v_cvt_sr_bf16_f32 ==
    # the oracle's BF16 formula (NaN, infinities and denormals included), and
    # the FP16 conversion built on v_cvt_sr_f16_f32 == its integer form, so
    # stochastic-rounding stores stay bit-exact against the oracle
    import ctypes as C
    from lattibol_amd import _lib
    bad = (C.c_uint64 * 2)(7, 7)
    _lib.check(_lib.lib.lb_check_sr_rounding(seed, bad), "lb_check_sr_rounding")
    assert (bad[0], bad[1]) == (0, 0)


@pytest.mark.gpu
@pytest.mark.parametrize("ci,co,H,W", [(16, 32, 16, 16), (32, 64, 8, 32), (96, 32, 16, 16), (384, 128, 16, 16),
                                       (64, 32, 8, 8), (16, 32, 8, 64), (96, 32, 16, 32), (384, 128, 8, 32),
                                       (64, 64, 24, 32), (32, 32, 128, 128), (256, 256, 16, 16), (16, 32, 32, 48),
                                       (64, 32, 16, 8)])
def test_gpu_conv3x3_wgrad_matches_reference(gpu, ci, co, H, W):
    from lattibol_amd import _lib
    rng = np.random.default_rng(ci * 7 + co + H)
    B = 2
    x = _bf16_round(rng.standard_normal((B, H, W, ci)))
    g = rng.standard_normal((B, H, W, co)).astype(np.float32)
    to_u16 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(torch.bfloat16).view(torch.int16).numpy()
    xd = _lib.DeviceArray.from_numpy(to_u16(x))
    gd = _lib.DeviceArray.from_numpy(g)
    out = {}
    for method in (0, 1, 2):
        dwd = _lib.DeviceArray((co, 9, ci), np.float32)
        rc = _lib.lib.lb_conv3x3_wgrad_bf16(method, gd.ptr, xd.ptr, dwd.ptr, B, H, W, ci, co, None)
        tile_ok = (W % 32 == 0 and H % 8 == 0) or (W % 16 == 0 and H % 16 == 0)
        if (method == 1 and W % 16) or (method == 2 and not tile_ok):
            assert rc == _lib.LB_ERR_UNSUPPORTED
            continue
        _lib.check(rc, "wgrad")
        out[method] = dwd.numpy()
    # float64 reference on the BF16-rounded operands: dW[co, tap, ci] =
    # sum_{n,h,w} g[n,h,w,co] x[n, h+dy, w+dx, ci] (zero padding)
    gb = _bf16_round(g).astype(np.float64)
    xp = np.zeros((B, H + 2, W + 2, ci))
    xp[:, 1:-1, 1:-1] = x
    ref = np.zeros((co, 9, ci))
    for tap in range(9):
        ky, kx = divmod(tap, 3)
        ref[:, tap] = np.einsum("nhwo,nhwi->oi", gb, xp[:, ky:ky + H, kx:kx + W])
    scale = np.abs(ref).max()
    for method, dw in out.items():
        np.testing.assert_allclose(dw, ref, rtol=1e-4, atol=1e-5 * scale, err_msg="method %d" % method)
    if 1 in out:
        assert np.array_equal(out[0], out[1])   # same operands, same summation order


# H = 32: every decoder level reads [upsampled low | skip] in place (LDS-tile
# convs and weight gradients); H = 16: the 8x8 level takes the materialised
# concatenation and the per-wavefront kernels.
@pytest.mark.gpu
@pytest.mark.parametrize("H", [32, 16])
def test_gpu_unet_forward_backward_match_oracle(gpu, H):
    from lattibol_amd import _lib
    from lattibol_amd.unet import UNet
    B = 2
    net = UNet(**SMALL, height=H, width=H, max_batch=B, seed=7)
    p, _ = _rand_params(**SMALL, seed=4)
    net.set_params(p)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((B, H, H, 3)).astype(np.float32)
    x[..., 0] = rng.uniform(size=(B, H, H)) < 0.2
    t = rng.standard_normal((B, H, H, 3)).astype(np.float32)
    m = (rng.uniform(size=(B, H, H)) > 0.1).astype(np.float32)
    xd, td, md = (_lib.DeviceArray.from_numpy(a) for a in (x, t, m))
    yd = _lib.DeviceArray((B, H, H, 3), np.float32)
    net.forward(xd, B, yd)
    loss = net.backward(td, md, B)
    y = yd.numpy()
    _, g = net.params(grads=True)
    params = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    yo = UO.forward(params, torch.from_numpy(x).double(), **SMALL)
    lo = UO.masked_mse(yo, torch.from_numpy(t).double(), torch.from_numpy(m).double())
    (go,) = torch.autograd.grad(lo, params)
    yo = yo.detach().numpy()
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    assert rel < 2e-2, rel
    assert abs(loss - lo.item()) < 3e-2 * lo.item()
    go = go.numpy()
    layout, _ = conv_layout(**SMALL)
    for kind, co, ci, w, b in layout:
        gw, ow = g[w:b], go[w:b]
        num = np.linalg.norm(gw - ow)
        den = np.linalg.norm(ow) + 1e-30
        cos = float(np.dot(gw, ow) / (np.linalg.norm(gw) * np.linalg.norm(ow) + 1e-30))
        assert num / den < 8e-2 and cos > 0.995, (kind, co, ci, num / den, cos)
    net.close()


@pytest.mark.gpu
def test_gpu_production_unet_gradient_matches_oracle(gpu):
    """The warm start's network as it trains: 4 levels, 32 base channels,
    128 x 128, He-scaled parameters, against the oracle that also rounds the
    pre-activation gradients to BF16 as the GPU stores them.  Measured: every
    weight and bias gradient within 2.7e-3 (relative L2) of the oracle's,
    cosine >= 0.999997; the remaining difference is FP32 summation order and
    BF16 activations that round the other way, so the emulation of the
    gradient rounding changes the figure in the fourth digit only."""
    from lattibol_amd import _lib
    from lattibol_amd.unet import UNet
    P = dict(levels=4, base=32, in_ch=3, out_ch=3)
    B, H = 2, 128
    net = UNet(**P, height=H, width=H, max_batch=B, seed=7)
    p, layout = _he_params(**P, seed=12)
    net.set_params(p)
    rng = np.random.default_rng(13)
    x = rng.standard_normal((B, H, H, 3)).astype(np.float32)
    x[..., 0] = rng.uniform(size=(B, H, H)) < 0.2
    t = rng.standard_normal((B, H, H, 3)).astype(np.float32)
    m = (rng.uniform(size=(B, H, H)) > 0.1).astype(np.float32)
    xd, td, md = (_lib.DeviceArray.from_numpy(a) for a in (x, t, m))
    yd = _lib.DeviceArray((B, H, H, 3), np.float32)
    net.forward(xd, B, yd)
    loss = net.backward(td, md, B)
    y = yd.numpy()
    _, g = net.params(grads=True)
    net.close()
    xt, tt, mt = (torch.from_numpy(a).double() for a in (x, t, m))
    res = {}
    for eg in (False, True):
        params = torch.tensor(p, dtype=torch.float64, requires_grad=True)
        yo = UO.forward(params, xt, **P, emulate_grad=eg)
        lo = UO.masked_mse(yo, tt, mt)
        (go,) = torch.autograd.grad(lo, params)
        res[eg] = (yo.detach().numpy(), lo.item(), go.numpy())
    yo, lo, go = res[True]
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print("output rel %.2e loss %.6f oracle %.6f" % (rel, loss, lo), flush=True)
    assert rel < 1e-2, rel
    assert abs(loss - lo) < 1e-2 * lo
    worst = {False: 0.0, True: 0.0}
    for i, (kind, co, ci, w, b) in enumerate(layout):
        for part, sl in (("w", slice(w, b)), ("b", slice(b, b + co))):
            gw = g[sl]
            errs = []
            for eg in (False, True):
                ow = res[eg][2][sl]
                errs.append(np.linalg.norm(gw - ow) / (np.linalg.norm(ow) + 1e-30))
                worst[eg] = max(worst[eg], errs[-1])
            cos = float(np.dot(gw, ow) / (np.linalg.norm(gw) * np.linalg.norm(ow) + 1e-30))
            print("%2d %s %s co %3d ci %3d  rel err plain %.2e  emulated %.2e  cos %.6f" % (
                i, kind, part, co, ci, errs[0], errs[1], cos), flush=True)
            assert errs[1] < 1e-2 and cos > 0.99995, (i, kind, part, errs, cos)
    print("worst rel err: plain oracle %.2e, with BF16 gradient emulation %.2e" % (worst[False], worst[True]),
          flush=True)


@pytest.mark.gpu
def test_gpu_unet_training_reduces_loss(gpu):
    from lattibol_amd import _lib
    from lattibol_amd.unet import UNet
    B, H = 4, 32
    net = UNet(**SMALL, height=H, width=H, max_batch=B, seed=11)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((B, H, H, 3)).astype(np.float32)
    t = np.tanh(x[..., ::-1]).astype(np.float32).copy()
    m = np.ones((B, H, H), np.float32)
    xd, td, md = (_lib.DeviceArray.from_numpy(a) for a in (x, t, m))
    losses = []
    for it in range(60):
        net.forward(xd, B)
        losses.append(net.backward(td, md, B))
        net.adam(lr=2e-3)
    net.sync()
    assert losses[-1] < 0.3 * losses[0], (losses[0], losses[-1])
    net.close()


@pytest.mark.gpu
def test_gpu_unet_backward_without_loss_readback(gpu):
    # the asynchronous form (no loss read-back, mask weight count reduced on
    # the device) gives the same gradients as the synchronous one; an
    # all-zero mask gives zero gradients and an error when the loss is asked for
    from lattibol_amd import _lib
    from lattibol_amd.unet import UNet
    B, H = 2, 32
    net = UNet(**SMALL, height=H, width=H, max_batch=B, seed=3)
    rng = np.random.default_rng(9)
    x = rng.standard_normal((B, H, H, 3)).astype(np.float32)
    t = rng.standard_normal((B, H, H, 3)).astype(np.float32)
    m = (rng.uniform(size=(B, H, H)) > 0.3).astype(np.float32)
    xd, td, md = (_lib.DeviceArray.from_numpy(a) for a in (x, t, m))
    net.forward(xd, B)
    assert net.backward(td, md, B, want_loss=False) is None
    _, g_async = net.params(grads=True)
    net.forward(xd, B)
    loss = net.backward(td, md, B)
    _, g_sync = net.params(grads=True)
    assert np.array_equal(g_async, g_sync)
    assert loss > 0
    zd = _lib.DeviceArray.from_numpy(np.zeros((B, H, H), np.float32))
    net.forward(xd, B)
    net.backward(td, zd, B, want_loss=False)
    _, g0 = net.params(grads=True)
    assert not np.any(g0)
    with pytest.raises(_lib.LatticeError):
        net.backward(td, zd, B)
    net.close()
