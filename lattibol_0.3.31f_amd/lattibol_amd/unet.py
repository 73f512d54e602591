"""U-Net surrogate of the steady channel flow (lb_unet_* in the C ABI).

Inputs per pixel: obstacle mask, standardised u_in, standardised tau
(3 channels, NHWC).  Outputs: standardised (ux/u_in, uy/u_in, (rho-1)/u_in).
Forward, backward and Adam run in liblattibol_amd.so on BF16 matrix cores.
"""
import ctypes as C

import numpy as np

from ._lib import DeviceArray, LbUnetDesc, check, lib

U_MEAN, U_STD = 0.05, 0.03          # u_in in [0.02, 0.08]
TAU_MEAN, TAU_STD = 0.71, 0.19      # tau in [0.52, 0.90]


def conv_layout(levels=4, base=32, in_ch=3, out_ch=3):
    """[(kind, co, ci, w_offset, b_offset)] in forward order, plus the head.

    Weight layout W[co][tap][ci] (tap = 3 ky + kx), first conv's ci padded to 16."""
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
    out = []
    off = 0
    for ci, co in convs:
        w = off
        off += co * 9 * ci
        out.append(("conv", co, ci, w, off))
        off += co
    w = off
    off += out_ch * C_[0]
    out.append(("head", out_ch, C_[0], w, off))
    off += out_ch
    return out, off


def make_inputs(solid, u_in, tau):
    """(B, H, W, 3) float32 network input from masks and per-sample u_in, tau."""
    solid = np.asarray(solid, np.float32)
    B = solid.shape[0]
    x = np.empty(solid.shape + (3,), np.float32)
    x[..., 0] = solid
    x[..., 1] = ((np.asarray(u_in, np.float64) - U_MEAN) / U_STD).reshape(B, 1, 1)
    x[..., 2] = ((np.asarray(tau, np.float64) - TAU_MEAN) / TAU_STD).reshape(B, 1, 1)
    return x


def make_targets(rho, ux, uy, u_in):
    """Physics-scaled targets (ux/u_in, uy/u_in, (rho-1)/u_in), (B, H, W, 3)."""
    u = np.asarray(u_in, np.float64).reshape(-1, 1, 1)
    t = np.stack([ux / u, uy / u, (rho - 1.0) / u], axis=-1)
    with np.errstate(over="ignore", invalid="ignore"):   # diverged channels: callers drop non-finite targets
        return t.astype(np.float32)


class UNet:
    def __init__(self, levels=4, base=32, in_ch=3, out_ch=3, height=128, width=128, max_batch=16, seed=0):
        d = LbUnetDesc(levels, base, in_ch, out_ch, height, width, max_batch)
        self.desc = d
        self._h = C.c_void_p()
        check(lib.lb_unet_create(C.byref(d), C.c_uint64(seed), C.byref(self._h)), "lb_unet_create")
        self.n_params = int(lib.lb_unet_param_count(self._h))
        self.layout, n = conv_layout(levels, base, in_ch, out_ch)
        assert n == self.n_params, (n, self.n_params)
        self.stats = np.array([0, 1, 0, 1, 0, 1], np.float32)   # target standardisation

    def close(self):
        if self._h:
            lib.lb_unet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self, grads=False):
        p = np.empty(self.n_params, np.float32)
        g = np.empty(self.n_params, np.float32) if grads else None
        check(lib.lb_unet_get_params(self._h, p.ctypes.data_as(C.c_void_p),
                                     g.ctypes.data_as(C.c_void_p) if grads else None), "lb_unet_get_params")
        return (p, g) if grads else p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        check(lib.lb_unet_set_params(self._h, p.ctypes.data_as(C.c_void_p)), "lb_unet_set_params")

    def forward(self, x_dev, B, y_dev=None):
        check(lib.lb_unet_forward(self._h, x_dev.ptr, int(B), y_dev.ptr if y_dev is not None else None),
              "lb_unet_forward")

    def backward(self, t_dev, mask_dev, B, want_loss=True):
        """Gradients of the masked MSE of the last forward; returns the loss
        (waits for the GPU) or, with want_loss=False, None (asynchronous)."""
        if not want_loss:
            check(lib.lb_unet_backward(self._h, t_dev.ptr, mask_dev.ptr, int(B), None), "lb_unet_backward")
            return None
        loss = C.c_double()
        check(lib.lb_unet_backward(self._h, t_dev.ptr, mask_dev.ptr, int(B), C.byref(loss)), "lb_unet_backward")
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib.lb_unet_adam(self._h, lr, beta1, beta2, eps), "lb_unet_adam")

    def sync(self):
        check(lib.lb_unet_sync(self._h), "lb_unet_sync")

    def to_fields(self, y_dev, B, uin_dev, fields_dev):
        s = np.ascontiguousarray(self.stats, np.float32)
        check(lib.lb_unet_to_fields(self._h, y_dev.ptr, int(B), uin_dev.ptr, s.ctypes.data_as(C.c_void_p),
                                    fields_dev.ptr), "lb_unet_to_fields")

    def predict(self, x, batch=None):
        """Host convenience: x (N,H,W,3) -> standardised y (N,H,W,out)."""
        N = x.shape[0]
        B = batch or self.desc.max_batch
        H, W = self.desc.height, self.desc.width
        xd = DeviceArray((B, H, W, self.desc.in_ch), np.float32)
        yd = DeviceArray((B, H, W, self.desc.out_ch), np.float32)
        out = np.empty((N, H, W, self.desc.out_ch), np.float32)
        for s in range(0, N, B):
            b = min(B, N - s)
            xb = np.zeros((B, H, W, self.desc.in_ch), np.float32)
            xb[:b] = x[s:s + b]
            xd.upload(xb)
            self.forward(xd, b, yd)
            out[s:s + b] = yd.numpy()[:b]
        return out
