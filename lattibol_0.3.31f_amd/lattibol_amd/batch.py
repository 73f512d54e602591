"""Batched ragged D2Q9 channels (lb_batch_* in the C ABI)."""
import ctypes as C

import numpy as np

from ._lib import LbBatchLattice, check, lib


class Batch:
    """Many independent D2Q9 channels stepped together on one GPU.

    Each lattice: the reference's Zou-He velocity inlet (u_in, 0) at x=0 and
    pressure outlet rho=1 at x=nx-1, half-way bounce-back walls at rows 0 and
    ny-1, solid obstacles with half-way bounce-back, FP32 deviation storage.
    run_to_steady keeps the whole loop on the device (one captured graph of
    many k-step chunks with on-device retirement and table compaction).
    """

    def __init__(self, specs):
        """specs: list of dicts with nx, ny, tau, u_in, solid (ny x nx bool)."""
        self.n = len(specs)
        arr = (LbBatchLattice * self.n)()
        masks = []
        off = 0
        self.shapes = []
        for i, s in enumerate(specs):
            nx, ny = int(s["nx"]), int(s["ny"])
            arr[i].nx, arr[i].ny = nx, ny
            arr[i].tau_inv = float(np.float32(1.0) / np.float32(s["tau"])) if "tau_inv" not in s else s["tau_inv"]
            arr[i].u_in = float(s["u_in"])
            arr[i].solid_offset = off
            m = np.asarray(s["solid"], dtype=np.uint8).reshape(ny, nx)
            masks.append(m.ravel())
            off += nx * ny
            self.shapes.append((ny, nx))
        self.masks = np.ascontiguousarray(np.concatenate(masks))
        self.specs = arr
        self._h = C.c_void_p()
        check(lib.lb_batch_create(self.n, arr, self.masks.ctypes.data_as(C.c_void_p), C.byref(self._h)),
              "lb_batch_create")
        self.total_nodes = int(lib.lb_batch_total_nodes(self._h))

    def close(self):
        if self._h:
            lib.lb_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_rest(self):
        check(lib.lb_batch_init_rest(self._h), "lb_batch_init_rest")

    def init_fields_device(self, ptr, mode=0):
        """Populations from packed device fields (rho, ux, uy per lattice).
        mode 0: equilibrium; 1: equilibrium + first-order non-equilibrium part."""
        if mode == 0:
            check(lib.lb_batch_init_fields_device(self._h, C.c_void_p(ptr)), "lb_batch_init_fields_device")
        else:
            check(lib.lb_batch_init_fields_device_ex(self._h, C.c_void_p(ptr), int(mode)),
                  "lb_batch_init_fields_device_ex")

    def step(self, n):
        check(lib.lb_batch_step(self._h, int(n)), "lb_batch_step")

    def run_to_steady(self, tol=1e-6, k=100, max_steps=200000):
        steps = np.zeros(self.n, np.int32)
        res = np.zeros(self.n, np.float64)
        check(lib.lb_batch_run_to_steady(self._h, float(tol), int(k), int(max_steps),
                                         steps.ctypes.data_as(C.c_void_p), res.ctypes.data_as(C.c_void_p)),
              "lb_batch_run_to_steady")
        return steps, res

    def macros(self, i):
        ny, nx = self.shapes[i]
        out = [np.empty((ny, nx), np.float32) for _ in range(3)]
        check(lib.lb_batch_get_macros(self._h, i, *[a.ctypes.data_as(C.c_void_p) for a in out]),
              "lb_batch_get_macros")
        return dict(zip(("rho", "ux", "uy"), out))

    def populations(self, i):
        ny, nx = self.shapes[i]
        out = np.empty((9, ny, nx), np.float32)
        check(lib.lb_batch_get_populations(self._h, i, out.ctypes.data_as(C.c_void_p)), "lb_batch_get_populations")
        return out
