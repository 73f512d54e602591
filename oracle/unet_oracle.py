"""CPU oracle of the U-Net surrogate (float64 torch on the CPU).

TEST INFRASTRUCTURE ONLY: the checker for lb_unet_* and lb_conv3x3_bf16.
It restates the network of lattibol_0.3.31f_amd/csrc/lb_unet.hip with the
same parameter layout and emulates the GPU's storage roundings (BF16
activations and weights, round-to-nearest-even) as straight-through
operations, so its autograd gradients are those of the network the GPU
evaluates.  Pinned by finite differences (tests/test_unet_oracle.py).
"""
import numpy as np
import torch
import torch.nn.functional as F


def bf16(t, emulate=True):
    if not emulate:
        return t
    r = t.to(torch.bfloat16).to(t.dtype)
    return t + (r - t).detach()


class _GradBF16(torch.autograd.Function):
    """Identity forward; the backward rounds the incoming gradient to BF16
    (nearest-even), as the GPU stores the pre-activation gradient it feeds
    to the weight- and input-gradient convolutions."""

    @staticmethod
    def forward(ctx, t):
        return t.view_as(t)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(g.dtype)


def conv_layout(levels, base, in_ch, out_ch):
    C_ = [base << l for l in range(levels)]
    convs = []
    for l in range(levels - 1):
        convs.append((16 if l == 0 else C_[l - 1], C_[l]))
        convs.append((C_[l], C_[l]))
    convs.append((C_[levels - 2], C_[levels - 1]))
    convs.append((C_[levels - 1], C_[levels - 1]))
    for l in range(levels - 2, -1, -1):
        convs.append((C_[l + 1] + C_[l], C_[l]))
        convs.append((C_[l], C_[l]))
    return convs, C_


def unpack(params, levels, base, in_ch, out_ch):
    """Flat parameter vector -> list of (W [co, ci, 3, 3], b) and head (Wh [o, c], bh)."""
    convs, C_ = conv_layout(levels, base, in_ch, out_ch)
    off = 0
    out = []
    for ci, co in convs:
        n = co * 9 * ci
        w = params[off:off + n].reshape(co, 3, 3, ci).permute(0, 3, 1, 2)
        off += n
        b = params[off:off + co]
        off += co
        out.append((w, b))
    n = out_ch * C_[0]
    wh = params[off:off + n].reshape(out_ch, C_[0])
    off += n
    bh = params[off:off + out_ch]
    return out, (wh, bh)


def forward(params, x, levels=4, base=32, in_ch=3, out_ch=3, emulate=True, emulate_grad=False):
    """params: float64 tensor (requires_grad ok); x: (B, H, W, in_ch) -> (B, H, W, out_ch).

    emulate_grad: also round each convolution's pre-activation gradient to
    BF16 on its way into the weight and input gradients (the bias gradient
    sums the unrounded values), as the GPU backward does."""
    convs, (wh, bh) = unpack(params, levels, base, in_ch, out_ch)
    B, H, W, _ = x.shape
    xin = torch.zeros(B, 16, H, W, dtype=params.dtype)
    xin[:, :in_ch] = x.permute(0, 3, 1, 2)
    cur = bf16(xin, emulate)
    k = 0
    skips = []

    def conv(t):
        nonlocal k
        w, b = convs[k]
        k += 1
        y = F.conv2d(t, bf16(w, emulate), padding=1)
        if emulate_grad:
            y = _GradBF16.apply(y)
        y = y + b.view(1, -1, 1, 1)
        return bf16(torch.relu(y), emulate)

    for l in range(levels):
        cur = conv(cur)
        cur = conv(cur)
        if l < levels - 1:
            skips.append(cur)
            cur = bf16(F.avg_pool2d(cur, 2), emulate)
    for l in range(levels - 2, -1, -1):
        up = F.interpolate(cur, scale_factor=2, mode="nearest")
        cur = torch.cat([up, skips[l]], dim=1)
        cur = conv(cur)
        cur = conv(cur)
    y = torch.einsum("bchw,oc->bhwo", cur, bf16(wh, emulate)) + bh
    return y


def masked_mse(y, t, mask):
    d = (y - t) * mask.unsqueeze(-1)
    cnt = mask.sum() * y.shape[-1]
    return (d * d).sum() / cnt


def conv3x3_ref(x_nhwc, w_cotapci, bias=None, relu=False):
    """float64 reference of lb_conv3x3_bf16 on given (already rounded) values."""
    co, _, ci = w_cotapci.shape
    w = torch.as_tensor(np.asarray(w_cotapci, np.float64)).reshape(co, 3, 3, ci).permute(0, 3, 1, 2)
    x = torch.as_tensor(np.asarray(x_nhwc, np.float64)).permute(0, 3, 1, 2)
    y = F.conv2d(x, w, padding=1)
    if bias is not None:
        y = y + torch.as_tensor(np.asarray(bias, np.float64)).view(1, -1, 1, 1)
    if relu:
        y = torch.relu(y)
    return y.permute(0, 2, 3, 1).numpy()
