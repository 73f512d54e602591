"""ctypes binding of liblattibol_amd.so (the C ABI in include/lattibol_amd.h).

The shared library is the product: every compute call goes through it.  If
it is missing the import fails loudly; there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "liblattibol_amd.so")

# ---------------------------------------------------------------------------
# Structs (field order = include/lattibol_amd.h)
# ---------------------------------------------------------------------------
FP = C.POINTER(C.c_float)
BP = C.POINTER(C.c_bool)


class LbDim3(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint)]


def _data_fields(names):
    return [(n, t) for n, t in names]


class DataD2Q9(C.Structure):
    """lattibol.h:29-63."""
    _fields_ = (
        [("ux", FP), ("uy", FP), ("rho", FP), ("solid", BP), ("solid_bound", BP)]
        + [("f%d" % i, FP) for i in range(9)]
        + [("border_f%d" % i, FP) for i in (1, 5, 8, 3, 6, 7)]
        + [("nb_border_f%d" % i, FP) for i in (1, 5, 8, 3, 6, 7)]
        + [("width", C.c_int), ("height", C.c_int), ("deviceCount", C.c_int), ("dev", C.c_int),
           ("plotBuffer", FP), ("valueCur", FP), ("stream", C.c_void_p),
           ("psi", FP), ("ueq_x", FP), ("ueq_y", FP)]
    )


class DataD3Q19(C.Structure):
    """lattibol.h:65-97."""
    _fields_ = (
        [("ux", FP), ("uy", FP), ("uz", FP), ("rho", FP), ("solid", BP), ("solid_bound", BP)]
        + [("f%d" % i, FP) for i in range(19)]
        + [("border_f%d" % i, FP) for i in (1, 7, 8, 9, 10, 2, 11, 12, 13, 14)]
        + [("nb_border_f%d" % i, FP) for i in (1, 7, 8, 9, 10, 2, 11, 12, 13, 14)]
        + [("width", C.c_int), ("height", C.c_int), ("depth", C.c_int), ("deviceCount", C.c_int),
           ("dev", C.c_int), ("plotBuffer", FP), ("valueCur", FP), ("stream", C.c_void_p)]
    )


class LbRefParams(C.Structure):
    _fields_ = [("deviceCount", C.c_int), ("width", C.c_int), ("height", C.c_int), ("depth", C.c_int),
                ("ux", C.c_float), ("uy", C.c_float), ("uz", C.c_float), ("rho", C.c_float),
                ("viscosity", C.c_float), ("Fx", C.c_float), ("Fy", C.c_float), ("Fz", C.c_float),
                ("G", C.c_float), ("Gads", C.c_float), ("zou_he", C.c_int),
                ("cursor_x", C.c_int), ("cursor_y", C.c_int), ("cursor_z", C.c_int)]


class LbConfig(C.Structure):
    _fields_ = [("lattice", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("storage", C.c_int), ("tau_inv", C.c_float), ("bc", C.c_int * 6),
                ("bc_u", (C.c_float * 3) * 6), ("bc_rho", C.c_float * 6), ("force", C.c_float * 3),
                ("sc_G", C.c_float), ("sc_rho0", C.c_float), ("sc_psi_solid", C.c_float),
                ("bb_full", C.c_int), ("zou_he", C.c_int), ("zh_ux_in", C.c_float),
                ("zh_rho_out", C.c_float), ("halo_z", C.c_int), ("block_x", C.c_int),
                ("block_y", C.c_int), ("xcd_remap", C.c_int), ("graph_chunk", C.c_int),
                ("rest_closure", C.c_int), ("mem_mode", C.c_int), ("layout", C.c_int), ("round_mode", C.c_int),
                ("node_words", C.c_int), ("time_block", C.c_int), ("sc_sweep", C.c_int), ("sc_zchunk", C.c_int)]


class LbBatchLattice(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("tau_inv", C.c_float), ("u_in", C.c_float),
                ("solid_offset", C.c_int64)]


class LbUnetDesc(C.Structure):
    _fields_ = [("levels", C.c_int), ("base", C.c_int), ("in_ch", C.c_int), ("out_ch", C.c_int),
                ("height", C.c_int), ("width", C.c_int), ("max_batch", C.c_int)]


class LbDeckParams(C.Structure):
    _fields_ = [("device_count", C.c_int), ("max_dim", C.c_int), ("update_steps", C.c_int),
                ("ux", C.c_float), ("uy", C.c_float), ("uz", C.c_float), ("rho", C.c_float),
                ("viscosity", C.c_float), ("tau_inv", C.c_float), ("fx", C.c_float), ("fy", C.c_float),
                ("fz", C.c_float), ("G", C.c_float), ("Gads", C.c_float), ("min_u", C.c_float),
                ("max_u", C.c_float), ("min_rho", C.c_float), ("max_rho", C.c_float)]


# ---------------------------------------------------------------------------
# Loading and prototypes
# ---------------------------------------------------------------------------
class LatticeError(RuntimeError):
    pass


if not os.path.exists(LIB_PATH):
    raise ImportError(
        "liblattibol_amd.so not found at %s: build it first (python -c 'import __graft_entry__ as g; g.build()')"
        % LIB_PATH)

lib = C.CDLL(LIB_PATH)

_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
_d3 = LbDim3

_PROTOS = {
    # error state
    "lb_last_error": (_i, []),
    "lb_error_string": (C.c_char_p, [_i]),
    "lb_device_count": (_i, []),
    "lb_device_alloc": (_i, [C.c_size_t, C.POINTER(_vp)]),
    "lb_device_free": (_i, [_vp]),
    "lb_memcpy_h2d": (_i, [_vp, _vp, C.c_size_t]),
    "lb_memcpy_d2h": (_i, [_vp, _vp, C.c_size_t]),
    "lb_device_synchronize": (_i, []),
    "lb_copy_device": (_i, [_vp, _vp, C.c_size_t, _i, FP]),
    # reference launchers
    "streaming_d2q9": (None, [C.POINTER(DataD2Q9), _d3, _d3]),
    "streaming_d3q19": (None, [C.POINTER(DataD3Q19), _d3, _d3]),
    "collision_d2q9": (None, [FP, C.POINTER(DataD2Q9)] + [_f] * 11 + [_d3, _d3] + [_f] * 4),
    "collision_d3q19": (None, [FP, C.POINTER(DataD3Q19)] + [_f] * 12 + [_d3, _d3]),
    "get_border_d2q9": (None, [C.POINTER(DataD2Q9), _d3, _d3]),
    "get_border_d3q19": (None, [C.POINTER(DataD3Q19), _d3, _d3]),
    "apply_border_d2q9": (None, [C.POINTER(DataD2Q9), _d3, _d3]),
    "apply_border_d3q19": (None, [C.POINTER(DataD3Q19), _d3, _d3]),
    "copy_border_devices_d2q9": (None, [FP] * 12 + [_i, C.POINTER(_vp)]),
    "copy_border_devices_d3q19": (None, [FP] * 20 + [_i, _i, C.POINTER(_vp)]),
    "read_value_2d": (None, [FP, C.POINTER(DataD2Q9), _i, _i, _i]),
    "read_value_3d": (None, [FP, C.POINTER(DataD3Q19), _i, _i, _i]),
    # reference driver
    "lb_refsim_create": (_i, [C.POINTER(LbRefParams), _vp, C.POINTER(_vp)]),
    "lb_refsim_destroy": (_i, [_vp]),
    "lb_refsim_run": (_i, [_vp, _i]),
    "lb_refsim_value_cur": (_i, [_vp, FP]),
    "lb_refsim_download": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lb_refsim_data_d2q9": (_i, [_vp, _i, C.POINTER(DataD2Q9)]),
    "lb_refsim_data_d3q19": (_i, [_vp, _i, C.POINTER(DataD3Q19)]),
    # fused solver
    "lb_create": (_i, [C.POINTER(LbConfig), C.POINTER(_vp)]),
    "lb_destroy": (_i, [_vp]),
    "lb_get_config": (_i, [_vp, C.POINTER(LbConfig)]),
    "lb_set_solid": (_i, [_vp, _vp]),
    "lb_init_equilibrium": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "lb_init_equilibrium_device": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "lb_init_uniform": (_i, [_vp, _f, _f, _f, _f]),
    "lb_set_populations": (_i, [_vp, _vp]),
    "lb_get_populations": (_i, [_vp, _vp]),
    "lb_step": (_i, [_vp, _i]),
    "lb_step_prepare": (_i, [_vp, _i]),
    "lb_sync": (_i, [_vp]),
    "lb_get_macros": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "lb_macros_device": (_i, [_vp, C.POINTER(_vp)]),
    "lb_reduce": (_i, [_vp, C.POINTER(C.c_double)]),
    "lb_residual": (_i, [_vp, C.POINTER(C.c_double)]),
    "lb_read_value": (_i, [_vp, _i, _i, _i, FP]),
    "lb_steps_done": (_i, [_vp, C.POINTER(C.c_longlong)]),
    "lb_time_step_kernel": (_i, [_vp, _i, FP]),
    "lb_halo_buffers": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    # batch
    "lb_batch_create": (_i, [_i, C.POINTER(LbBatchLattice), _vp, C.POINTER(_vp)]),
    "lb_batch_destroy": (_i, [_vp]),
    "lb_batch_init_rest": (_i, [_vp]),
    "lb_batch_init_fields_device": (_i, [_vp, _vp]),
    "lb_batch_init_fields_device_ex": (_i, [_vp, _vp, _i]),
    "lb_batch_step": (_i, [_vp, _i]),
    "lb_batch_run_to_steady": (_i, [_vp, C.c_double, _i, _i, _vp, _vp]),
    "lb_batch_get_macros": (_i, [_vp, _i, _vp, _vp, _vp]),
    "lb_batch_get_populations": (_i, [_vp, _i, _vp]),
    "lb_batch_total_nodes": (C.c_int64, [_vp]),
    # U-Net surrogate
    "lb_unet_create": (_i, [C.POINTER(LbUnetDesc), C.c_uint64, C.POINTER(_vp)]),
    "lb_unet_destroy": (_i, [_vp]),
    "lb_unet_param_count": (C.c_int64, [_vp]),
    "lb_unet_get_params": (_i, [_vp, _vp, _vp]),
    "lb_unet_set_params": (_i, [_vp, _vp]),
    "lb_unet_forward": (_i, [_vp, _vp, _i, _vp]),
    "lb_unet_backward": (_i, [_vp, _vp, _vp, _i, C.POINTER(C.c_double)]),
    "lb_unet_adam": (_i, [_vp, _f, _f, _f, _f]),
    "lb_unet_sync": (_i, [_vp]),
    "lb_unet_to_fields": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "lb_conv3x3_bf16": (_i, [_i, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "lb_conv3x3_bf16_noalloc": (_i, [_i, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "lb_conv3x3_wgrad_bf16": (_i, [_i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "lb_check_bf16_rounding": (_i, [C.POINTER(C.c_uint64)]),
    "lb_check_sr_rounding": (_i, [C.c_uint32, C.POINTER(C.c_uint64)]),
    "lb_deck_read_params": (_i, [C.c_char_p, C.POINTER(LbDeckParams)]),
    "lb_deck_read_bmp": (_i, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _vp]),
    "lb_deck_voxelize_stl": (_i, [C.c_char_p, _i, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _vp]),
    "lb_write_pvti": (_i, [C.c_char_p, _i, _i, _i, _i, _i]),
    "lb_write_vti": (_i, [C.c_char_p, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
}

EXPORTED = sorted(_PROTOS)
OPTIONAL = set()

for _name, (_res, _args) in _PROTOS.items():
    _fn = getattr(lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def bind_optional(name, res, args):
    """Bind a symbol added by a later component, if this build exports it."""
    fn = getattr(lib, name, None)
    if fn is None:
        return None
    fn.restype = res
    fn.argtypes = args
    _PROTOS[name] = (res, args)
    return fn


# error codes (include/lattibol_amd.h)
LB_OK, LB_ERR_INVALID, LB_ERR_SHAPE, LB_ERR_UNSUPPORTED, LB_ERR_NO_DEVICE, LB_ERR_NOT_READY = 0, -1, -2, -3, -4, -5
LB_ERR_SCRATCH = -6


def check(rc, what=""):
    """Raise LatticeError for a nonzero return code of an lb_* call."""
    if rc:
        msg = lib.lb_error_string(int(rc)).decode()
        raise LatticeError("%s failed: %s (code %d)" % (what or "lattibol call", msg, rc))
    return rc


def check_sticky(what=""):
    """Raise on the sticky error set by the void reference launchers."""
    rc = lib.lb_last_error()
    if rc:
        check(rc, what)


def device_count():
    return int(lib.lb_device_count())


class DeviceArray:
    """A device allocation with numpy upload/download (plumbing for ctypes hosts)."""

    def __init__(self, shape, dtype):
        import numpy as np
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        check(lib.lb_device_alloc(self.nbytes, C.byref(self.ptr)), "lb_device_alloc")

    @classmethod
    def from_numpy(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib.lb_memcpy_h2d(self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes), "lb_memcpy_h2d")

    def numpy(self):
        import numpy as np
        out = np.empty(self.shape, self.dtype)
        check(lib.lb_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes), "lb_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib.lb_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
