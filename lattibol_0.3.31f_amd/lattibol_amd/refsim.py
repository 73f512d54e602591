"""The reference host driver (Simulation, SimulationD2Q9, SimulationD3Q19 in
src/simulation*.cu) as a thin wrapper over lb_refsim_* in the C ABI.

Same construction inputs as the reference's Input deck (input.cpp:51-97:
deviceCount, velocities, density, viscosity, forces, Shan-Chen G / Gads) plus
the solid mask, and the same runLBM() step loop (stream, halo exchange over
the ring of x-slab subdomains, collide, probe).
"""
import ctypes as C

import numpy as np

from ._lib import LbRefParams, check, lib, DataD2Q9, DataD3Q19


class Simulation:
    def __init__(self, solids, deviceCount=1, ux=0.0, uy=0.0, uz=0.0, rho=1.0, viscosity=1.0 / 6.0,
                 Fx=0.0, Fy=0.0, Fz=0.0, G=0.0, Gads=0.0, cursor=(0, 0, 0), zou_he=True):
        """zou_he: the reference's always-on Zou-He inlet (x = 0) / outlet
        (x = W-1) (collision.cu:211-262); False makes x periodic."""
        s = np.ascontiguousarray(np.asarray(solids, dtype=np.uint8))
        if s.ndim == 2:
            s = s[None]
        self.depth, self.height, self.width = s.shape
        self.is3d = self.depth > 1
        self.Q = 19 if self.is3d else 9
        self.deviceCount = int(deviceCount)
        p = LbRefParams()
        p.deviceCount = self.deviceCount
        p.width, p.height, p.depth = self.width, self.height, self.depth
        p.ux, p.uy, p.uz, p.rho, p.viscosity = ux, uy, uz, rho, viscosity
        p.Fx, p.Fy, p.Fz, p.G, p.Gads = Fx, Fy, Fz, G, Gads
        p.zou_he = 1 if zou_he else 0
        p.cursor_x, p.cursor_y, p.cursor_z = cursor
        self.params = p
        self.steps = 0
        self._h = C.c_void_p()
        check(lib.lb_refsim_create(C.byref(p), s.ctypes.data_as(C.c_void_p), C.byref(self._h)), "lb_refsim_create")
        self.sub_width = self.width // self.deviceCount

    def close(self):
        if self._h:
            lib.lb_refsim_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def runLBM(self, updateSteps):
        check(lib.lb_refsim_run(self._h, int(updateSteps)), "lb_refsim_run")
        self.steps += int(updateSteps)

    def getSteps(self):
        return self.steps

    def valueCur(self):
        out = (C.c_float * 4)()
        check(lib.lb_refsim_value_cur(self._h, out), "lb_refsim_value_cur")
        return np.array(out[:], np.float32)

    def download(self, dev=0):
        N = self.sub_width * self.height * self.depth
        shp = (self.depth, self.height, self.sub_width)
        f = np.empty((self.Q,) + shp, np.float32)
        ux, uy, uz, rho = (np.empty(shp, np.float32) for _ in range(4))
        plot = np.empty((8 if not self.is3d else 4) * N, np.float32)
        args = [a.ctypes.data_as(C.c_void_p) for a in (f, ux, uy, uz, rho, plot)]
        check(lib.lb_refsim_download(self._h, dev, *args), "lb_refsim_download")
        return dict(f=f, ux=ux, uy=uy, uz=uz, rho=rho, plot=plot)

    def upload_populations(self, f, dev=0):
        """Overwrite subdomain dev's populations (Q x N float32, f0 plane
        first) through the device pointers of its Data struct, e.g. to start
        from a droplet instead of the reference's uniform density."""
        f = np.ascontiguousarray(np.asarray(f, np.float32).reshape(self.Q, -1))
        d = self.data(dev)
        for i in range(self.Q):
            ptr = C.cast(getattr(d, "f%d" % i), C.c_void_p)
            check(lib.lb_memcpy_h2d(ptr, f[i].ctypes.data_as(C.c_void_p), f[i].nbytes), "lb_memcpy_h2d")

    def data(self, dev=0):
        if self.is3d:
            d = DataD3Q19()
            check(lib.lb_refsim_data_d3q19(self._h, dev, C.byref(d)), "lb_refsim_data_d3q19")
        else:
            d = DataD2Q9()
            check(lib.lb_refsim_data_d2q9(self._h, dev, C.byref(d)), "lb_refsim_data_d2q9")
        return d


def SimulationD2Q9(solids, **kw):
    s = np.asarray(solids)
    assert s.ndim == 2
    return Simulation(s, **kw)


def SimulationD3Q19(solids, **kw):
    s = np.asarray(solids)
    assert s.ndim == 3
    return Simulation(s, **kw)
