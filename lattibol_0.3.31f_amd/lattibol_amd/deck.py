"""Input decks and VTK output (lb_deck_* / lb_write_* in the C ABI).

A deck is a parameter file `name.dat` plus a solids file of the same name:
`name.bmp` for 2D decks (maxDim <= 1) or `name.stl` for 3D decks, the pairing
MainWindow::loadFile makes (mainwindow.cpp:236-275).  `load()` returns the
parameters and the solid mask; `simulation()` builds the reference driver
from them, `write_vtk()` writes the reference's .pvti/.vti files.
"""
import ctypes as C
import os

import numpy as np

from ._lib import LbDeckParams, check, lib

PARAM_FIELDS = ("device_count", "max_dim", "update_steps", "ux", "uy", "uz", "rho", "viscosity", "tau_inv", "fx",
                "fy", "fz", "G", "Gads", "min_u", "max_u", "min_rho", "max_rho")


def read_params(dat_path):
    p = LbDeckParams()
    check(lib.lb_deck_read_params(os.fsencode(dat_path), C.byref(p)), "lb_deck_read_params(%s)" % dat_path)
    return {k: getattr(p, k) for k in PARAM_FIELDS}


def read_bmp(bmp_path):
    """(H, W) uint8 mask, 1 = solid (black pixel), row 0 = top of the image."""
    w, h = C.c_int(), C.c_int()
    path = os.fsencode(bmp_path)
    check(lib.lb_deck_read_bmp(path, C.byref(w), C.byref(h), None), "lb_deck_read_bmp(%s)" % bmp_path)
    out = np.zeros((h.value, w.value), np.uint8)
    check(lib.lb_deck_read_bmp(path, C.byref(w), C.byref(h), out.ctypes.data_as(C.c_void_p)), "lb_deck_read_bmp")
    return out


def voxelize_stl(stl_path, max_dim):
    """(D, H, W) uint8 mask, the mesh's longest extent spanning max_dim voxels."""
    w, h, d = C.c_int(), C.c_int(), C.c_int()
    path = os.fsencode(stl_path)
    check(lib.lb_deck_voxelize_stl(path, int(max_dim), C.byref(w), C.byref(h), C.byref(d), None),
          "lb_deck_voxelize_stl(%s)" % stl_path)
    out = np.zeros((d.value, h.value, w.value), np.uint8)
    check(lib.lb_deck_voxelize_stl(path, int(max_dim), C.byref(w), C.byref(h), C.byref(d),
                                   out.ctypes.data_as(C.c_void_p)), "lb_deck_voxelize_stl")
    return out


def load(dat_path):
    """Parameters and solids of a deck: {'params': {...}, 'solids': mask}."""
    params = read_params(dat_path)
    stem = os.path.splitext(dat_path)[0]
    if params["max_dim"] > 1:
        solids = voxelize_stl(stem + ".stl", params["max_dim"])
    else:
        solids = read_bmp(stem + ".bmp")
    return dict(params=params, solids=solids, path=dat_path)


def simulation(deck, cursor=(0, 0, 0), **overrides):
    """The reference driver (refsim.Simulation) configured from a deck;
    keyword overrides (e.g. zou_he=False, Gads=...) replace deck values."""
    from .refsim import Simulation
    p = deck["params"]
    kw = dict(deviceCount=max(1, p["device_count"]), ux=p["ux"], uy=p["uy"], uz=p["uz"], rho=p["rho"],
              viscosity=p["viscosity"], Fx=p["fx"], Fy=p["fy"], Fz=p["fz"], G=p["G"], Gads=p["Gads"], cursor=cursor)
    kw.update(overrides)
    return Simulation(deck["solids"], **kw)


def write_vtk(directory, vtk_id, fields, device_count=1):
    """Reference-layout VTK output of (rho, ux, uy[, uz]) fields shaped (D, H, W)
    or (H, W): one .vti piece per x-slab of width W / device_count and the
    .pvti index."""
    rho = np.asarray(fields["rho"], np.float32)
    shape = rho.shape if rho.ndim == 3 else (1,) + rho.shape
    D, H, W = shape
    if W % device_count:
        raise ValueError("width %d is not divisible by %d devices" % (W, device_count))
    w = W // device_count
    arrs = {k: np.asarray(fields.get(k, np.zeros(shape)), np.float32).reshape(shape) for k in ("rho", "ux", "uy", "uz")}
    d = os.fsencode(directory)
    check(lib.lb_write_pvti(d, int(vtk_id), device_count, w, H, D), "lb_write_pvti")
    for dev in range(device_count):
        piece = {k: np.ascontiguousarray(a[:, :, dev * w:(dev + 1) * w]) for k, a in arrs.items()}
        ptr = {k: v.ctypes.data_as(C.c_void_p) for k, v in piece.items()}
        check(lib.lb_write_vti(d, int(vtk_id), dev, w, H, D, ptr["ux"], ptr["uy"], ptr["uz"], ptr["rho"]),
              "lb_write_vti")
