"""Host-side mirror of the fused pull-scheme solver (lb_* in the C ABI).

A `Solver` owns one lattice on the GPU: two population buffers in the
chosen storage format, node flags, macro planes and (for the z-slab mode)
halo planes.  Everything numeric runs in liblattibol_amd.so.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import LbConfig, check, lib

BC = {"periodic": 0, "wall": 1, "moving_wall": 2, "inlet": 3, "outflow": 4}
STORAGE = {"f32": 0, "f16": 1, "bf16": 2, "f64": 3, "f32s": 4}
RAW_DTYPE = {0: np.float32, 1: np.uint16, 2: np.uint16, 3: np.float64, 4: np.float32}
FACES = ("-x", "+x", "-y", "+y", "-z", "+z")
LAYOUT = {"auto": -1, "soa": 0, "rows": 1}
# 16-bit rounding: "default" = stochastic on Shan-Chen lattices, nearest-even
# elsewhere; "nearest" forces nearest-even (for comparisons)
ROUND = {"default": 0, "stochastic": 1, "nearest": 2}

# Lattice weights as float32 (simulation_d2q9.cu:56-58, simulation_d3q19.cu:54-56).
W9 = np.array([16 / 36] + [4 / 36] * 4 + [1 / 36] * 4, dtype=np.float64)
W19 = np.array([12 / 36] + [2 / 36] * 6 + [1 / 36] * 12, dtype=np.float64)


def make_config(lattice, shape, tau=None, tau_inv=None, storage="f32", bc=None, bc_u=None, bc_rho=None,
                force=None, sc_G=0.0, sc_rho0=1.0, sc_psi_solid=0.0, bb_full=False, zou_he=False,
                zh_ux_in=0.0, zh_rho_out=1.0, halo_z=False, block_x=0, block_y=0, xcd_remap=-1, graph_chunk=0,
                rest_closure=None, mem_mode=-1, layout=-1, round_mode="default", node_words=-1, time_block=-1,
                sc_sweep=-1, sc_zchunk=0):
    """Plain-dict description of a lattice; the oracle binding reads the same dict.

    rest_closure (default: on unless bb_full) sets feq_0 = rho - sum_{i>0} feq_i so
    that collisions balance mass exactly up to rounding; off reproduces the
    reference's equilibrium for every direction (needed for its full-way
    bounce-back decks)."""
    if rest_closure is None:
        rest_closure = not bb_full
    nx, ny = shape[0], shape[1]
    nz = shape[2] if len(shape) > 2 else 1
    if tau_inv is None:
        tau_inv = float(np.float32(1.0) / np.float32(tau))
    bcs = [0] * 6
    for k, v in (bc or {}).items():
        bcs[FACES.index(k) if isinstance(k, str) else k] = BC[v] if isinstance(v, str) else int(v)
    us = [[0.0, 0.0, 0.0] for _ in range(6)]
    for k, v in (bc_u or {}).items():
        us[FACES.index(k) if isinstance(k, str) else k] = [float(a) for a in (list(v) + [0, 0, 0])[:3]]
    rs = [1.0] * 6
    for k, v in (bc_rho or {}).items():
        rs[FACES.index(k) if isinstance(k, str) else k] = float(v)
    return dict(lattice=int(lattice), nx=int(nx), ny=int(ny), nz=int(nz),
                storage=STORAGE[storage] if isinstance(storage, str) else int(storage),
                tau_inv=float(tau_inv), bc=bcs, bc_u=us, bc_rho=rs,
                force=[float(a) for a in (list(force or []) + [0, 0, 0])[:3]],
                sc_G=float(sc_G), sc_rho0=float(sc_rho0), sc_psi_solid=float(sc_psi_solid),
                bb_full=int(bool(bb_full)), zou_he=int(bool(zou_he)), zh_ux_in=float(zh_ux_in),
                zh_rho_out=float(zh_rho_out), halo_z=int(bool(halo_z)), block_x=int(block_x), block_y=int(block_y),
                xcd_remap=int(xcd_remap), graph_chunk=int(graph_chunk), rest_closure=int(bool(rest_closure)),
                mem_mode=int(mem_mode), layout=LAYOUT.get(layout, layout),
                round_mode=ROUND.get(round_mode, round_mode), node_words=int(node_words),
                time_block=int(time_block), sc_sweep=int(sc_sweep), sc_zchunk=int(sc_zchunk))


def to_lbconfig(d):
    c = LbConfig()
    defaults = dict(block_x=0, block_y=0, xcd_remap=-1, graph_chunk=0, mem_mode=-1, layout=-1, round_mode=0, node_words=-1,
                    time_block=-1, sc_sweep=-1, sc_zchunk=0)
    for k in ("lattice", "nx", "ny", "nz", "storage", "tau_inv", "sc_G", "sc_rho0", "sc_psi_solid", "bb_full",
              "zou_he", "zh_ux_in", "zh_rho_out", "halo_z", "block_x", "block_y", "xcd_remap", "graph_chunk", "rest_closure",
              "mem_mode", "layout", "round_mode", "node_words", "time_block", "sc_sweep", "sc_zchunk"):
        setattr(c, k, d[k] if k in d else defaults[k])
    for f in range(6):
        c.bc[f] = d["bc"][f]
        c.bc_rho[f] = d["bc_rho"][f]
        for a in range(3):
            c.bc_u[f][a] = d["bc_u"][f][a]
    for a in range(3):
        c.force[a] = d["force"][a]
    return c


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Solver:
    """One lattice on the GPU, stepped by the fused pull kernel."""

    def __init__(self, cfg, solid=None):
        if isinstance(cfg, dict):
            self.cfg = dict(cfg)
        else:
            raise TypeError("cfg must come from make_config()")
        self.Q = self.cfg["lattice"]
        self.shape = (self.cfg["nz"], self.cfg["ny"], self.cfg["nx"])   # array order z, y, x
        self.N = int(np.prod(self.shape))
        self._h = C.c_void_p()
        check(lib.lb_create(C.byref(to_lbconfig(self.cfg)), C.byref(self._h)), "lb_create")
        if solid is not None:
            self.set_solid(solid)

    # -- lifetime --------------------------------------------------------
    def close(self):
        if self._h:
            lib.lb_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- set-up ----------------------------------------------------------
    @property
    def launch_config(self):
        c = LbConfig()
        check(lib.lb_get_config(self._h, C.byref(c)), "lb_get_config")
        return dict(block_x=c.block_x, block_y=c.block_y, xcd_remap=c.xcd_remap, mem_mode=c.mem_mode, layout=c.layout,
                    node_words=c.node_words, time_block=c.time_block,
                    graph_chunk=c.graph_chunk, sc_sweep=c.sc_sweep, sc_zchunk=c.sc_zchunk,
                    round_mode={1: "stochastic", 2: "nearest"}.get(c.round_mode, "default"))

    def set_solid(self, solid):
        s = np.ascontiguousarray(np.asarray(solid, dtype=np.uint8).reshape(self.shape))
        check(lib.lb_set_solid(self._h, _ptr(s)), "lb_set_solid")

    def init_equilibrium(self, rho, ux, uy, uz=None):
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float32), self.shape))
                if a is not None else None for a in (rho, ux, uy, uz)]
        check(lib.lb_init_equilibrium(self._h, *[_ptr(a) for a in arrs]), "lb_init_equilibrium")

    def init_equilibrium_device(self, rho_ptr, ux_ptr, uy_ptr, uz_ptr=None):
        check(lib.lb_init_equilibrium_device(self._h, rho_ptr, ux_ptr, uy_ptr, uz_ptr), "lb_init_equilibrium_device")

    def init_uniform(self, rho=1.0, ux=0.0, uy=0.0, uz=0.0):
        check(lib.lb_init_uniform(self._h, rho, ux, uy, uz), "lb_init_uniform")

    def set_populations(self, raw):
        a = np.ascontiguousarray(np.asarray(raw, dtype=RAW_DTYPE[self.cfg["storage"]]))
        if a.size != self.Q * self.N:
            raise ValueError("expected %d populations" % (self.Q * self.N))
        check(lib.lb_set_populations(self._h, _ptr(a)), "lb_set_populations")

    # -- stepping --------------------------------------------------------
    def step(self, n=1):
        check(lib.lb_step(self._h, int(n)), "lb_step")

    def prepare(self, n):
        """Instantiate the graphs step(n) will launch, without stepping."""
        check(lib.lb_step_prepare(self._h, int(n)), "lb_step_prepare")

    def sync(self):
        check(lib.lb_sync(self._h), "lb_sync")

    @property
    def steps_done(self):
        v = C.c_longlong()
        check(lib.lb_steps_done(self._h, C.byref(v)), "lb_steps_done")
        return v.value

    def time_step_kernel(self, n=20):
        ms = C.c_float()
        check(lib.lb_time_step_kernel(self._h, int(n), C.byref(ms)), "lb_time_step_kernel")
        return ms.value

    # -- read-back -------------------------------------------------------
    def populations(self):
        out = np.empty((self.Q,) + self.shape, dtype=RAW_DTYPE[self.cfg["storage"]])
        check(lib.lb_get_populations(self._h, _ptr(out)), "lb_get_populations")
        return out

    def macros(self):
        out = {k: np.empty(self.shape, np.float32) for k in ("rho", "ux", "uy", "uz")}
        check(lib.lb_get_macros(self._h, _ptr(out["rho"]), _ptr(out["ux"]), _ptr(out["uy"]), _ptr(out["uz"])),
              "lb_get_macros")
        return out

    def macros_device_ptr(self):
        p = C.c_void_p()
        check(lib.lb_macros_device(self._h, C.byref(p)), "lb_macros_device")
        return p.value

    def reduce(self):
        out = (C.c_double * 4)()
        check(lib.lb_reduce(self._h, out), "lb_reduce")
        return np.array(out[:], dtype=np.float64)

    def residual(self):
        v = C.c_double()
        check(lib.lb_residual(self._h, C.byref(v)), "lb_residual")
        return v.value

    def read_value(self, x, y, z=0):
        out = (C.c_float * 4)()
        check(lib.lb_read_value(self._h, x, y, z, out), "lb_read_value")
        return np.array(out[:], dtype=np.float32)

    def halo_buffers(self):
        ps = [C.c_void_p() for _ in range(4)]
        check(lib.lb_halo_buffers(self._h, *[C.byref(p) for p in ps]), "lb_halo_buffers")
        return dict(zip(("send_lo", "send_hi", "recv_lo", "recv_hi"), [p.value for p in ps]))


def decode_populations(raw, storage, lattice):
    """Stored populations -> float64 (adds w_i back for the deviation storages)."""
    st = STORAGE[storage] if isinstance(storage, str) else storage
    w = (W9 if lattice == 9 else W19).astype(np.float32)
    if st == 0:
        return raw.astype(np.float64)
    if st == 3:
        return raw.astype(np.float64)
    if st == 4:
        v = raw.astype(np.float32)
    elif st == 1:
        v = raw.view(np.float16).astype(np.float32)
    else:
        v = (raw.astype(np.uint32) << 16).view(np.float32)
    shp = (-1,) + (1,) * (raw.ndim - 1)
    return v.astype(np.float64) + w.astype(np.float64).reshape(shp)
