"""ctypes binding of the CPU oracle (liblb_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblb_oracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


if not os.path.exists(LIB_PATH):
    build()

lib = C.CDLL(LIB_PATH)


class OrcCfg(C.Structure):
    _fields_ = [("lattice", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("storage", C.c_int), ("tau_inv", C.c_float), ("bc", C.c_int * 6),
                ("bc_u", (C.c_float * 3) * 6), ("bc_rho", C.c_float * 6), ("force", C.c_float * 3),
                ("sc_G", C.c_float), ("sc_rho0", C.c_float), ("sc_psi_solid", C.c_float),
                ("bb_full", C.c_int), ("zou_he", C.c_int), ("zh_ux_in", C.c_float),
                ("zh_rho_out", C.c_float), ("rest_closure", C.c_int), ("round_mode", C.c_int),
                ("step", C.c_uint32)]

    @classmethod
    def from_dict(cls, d):
        c = cls()
        for k in ("lattice", "nx", "ny", "nz", "storage", "tau_inv", "sc_G", "sc_rho0", "sc_psi_solid",
                  "bb_full", "zou_he", "zh_ux_in", "zh_rho_out", "rest_closure"):
            setattr(c, k, d[k])
        # the library's round_mode: 0 default (stochastic for Shan-Chen with
        # 16-bit storage), 1 stochastic, 2 nearest-even; the oracle's: 0 / 1
        rm = int(d.get("round_mode", 0))
        sixteen = d["storage"] in (1, 2)
        c.round_mode = int(sixteen and (rm == 1 or (rm == 0 and d["sc_G"] != 0.0)))
        c.step = 0
        for f in range(6):
            c.bc[f] = d["bc"][f]
            c.bc_rho[f] = d["bc_rho"][f]
            for a in range(3):
                c.bc_u[f][a] = d["bc_u"][f][a]
        for a in range(3):
            c.force[a] = d["force"][a]
        return c


vp = C.c_void_p
i_ = C.c_int
f_ = C.c_float
_PROTOS = {
    "orc_exp": (C.c_double, [C.c_double]),
    "orc_expf": (C.c_float, [C.c_float]),
    "orc_psi_f32": (C.c_float, [C.c_float, C.c_float]),
    "orc_psi_f64": (C.c_double, [C.c_double, C.c_float]),
    "orc_splitmix64_next": (C.c_uint64, [C.POINTER(C.c_uint64)]),
    "orc_uniform_noise": (None, [C.c_uint64, C.c_int64, C.c_double, C.c_double, vp]),
    "orc_f32_to_f16": (C.c_uint16, [f_]),
    "orc_f16_to_f32": (f_, [C.c_uint16]),
    "orc_f32_to_bf16": (C.c_uint16, [f_]),
    "orc_sr_bits": (C.c_uint32, [C.c_uint32, C.c_uint32, i_]),
    "orc_f32_to_f16_sr": (C.c_uint16, [f_, C.c_uint32]),
    "orc_f32_to_bf16_sr": (C.c_uint16, [f_, C.c_uint32]),
    "orc_bf16_to_f32": (f_, [C.c_uint16]),
    "orc_f32_to_f16_array": (None, [vp, vp, C.c_int64]),
    "orc_f32_to_bf16_array": (None, [vp, vp, C.c_int64]),
    "orc_set_threads": (None, [i_]),
    "orc_get_threads": (i_, []),
    "orc_ref_streaming_d2q9": (None, [vp, vp, i_, i_]),
    "orc_ref_streaming_d3q19": (None, [vp, vp, i_, i_, i_]),
    "orc_ref_collision_d2q9": (None, [vp] * 9 + [i_] * 4 + [f_] * 11 + [i_]),
    "orc_ref_collision_d3q19": (None, [vp] * 7 + [i_] * 5 + [f_] * 8 + [i_]),
    "orc_ref_get_border_d2q9": (None, [vp, vp, i_, i_]),
    "orc_ref_apply_border_d2q9": (None, [vp, vp, i_, i_]),
    "orc_ref_get_border_d3q19": (None, [vp, vp, i_, i_, i_]),
    "orc_ref_apply_border_d3q19": (None, [vp, vp, i_, i_, i_]),
    "orc_ref_read_value": (None, [vp, i_, i_, i_, i_, i_, vp]),
    "orc_ref_solid_bound_d2q9": (None, [vp, vp, i_, i_]),
    "orc_ref_solid_bound_d3q19": (None, [vp, vp, i_, i_, i_]),
    "orc_ref_init_d2q9": (i_, [vp, i_, i_, i_] + [f_] * 8 + [vp] * 9),
    "orc_ref_init_d3q19": (i_, [vp, i_, i_, i_, i_] + [f_] * 4 + [vp] * 7),
    "orc_ref_run_d2q9": (None, [i_, i_, i_] + [vp] * 10 + [f_] * 8 + [i_, i_, i_, vp, i_]),
    "orc_ref_run_d3q19": (None, [i_, i_, i_, i_] + [vp] * 8 + [f_] * 5 + [i_, i_, i_, i_, vp, i_]),
    "orc_pull_init_eq": (None, [C.POINTER(OrcCfg), vp, vp, vp, vp, vp]),
    "orc_pull_step": (None, [C.POINTER(OrcCfg), vp, vp, vp, vp]),
    "orc_pull_run": (None, [C.POINTER(OrcCfg), vp, vp, vp, i_, vp]),
    "orc_pull_decode": (None, [C.POINTER(OrcCfg), vp, vp]),
    "orc_pull_reduce": (None, [C.POINTER(OrcCfg), vp, vp, vp]),
    "orc_residual": (C.c_double, [i_, C.c_int64, vp, vp, vp]),
}
for _n, (_r, _a) in _PROTOS.items():
    _fn = getattr(lib, _n)
    _fn.restype = _r
    _fn.argtypes = _a

RAW = {0: np.float32, 1: np.uint16, 2: np.uint16, 3: np.float64, 4: np.float32}
W9F = np.array([16 / 36] + [4 / 36] * 4 + [1 / 36] * 4, dtype=np.float32)
W19F = np.array([12 / 36] + [2 / 36] * 6 + [1 / 36] * 12, dtype=np.float32)


def P(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def set_threads(n):
    lib.orc_set_threads(int(n))


# ---------------------------------------------------------------------------
# Pull-scheme specification
# ---------------------------------------------------------------------------
class PullOracle:
    """CPU oracle of the fused solver for one configuration dict (make_config)."""

    def __init__(self, cfg, solid=None):
        self.cfg = dict(cfg)
        self.c = OrcCfg.from_dict(self.cfg)
        self.Q = self.cfg["lattice"]
        self.shape = (self.cfg["nz"], self.cfg["ny"], self.cfg["nx"])
        self.N = int(np.prod(self.shape))
        dt = RAW[self.cfg["storage"]]
        self.A = np.zeros((self.Q,) + self.shape, dt)
        self.B = np.zeros_like(self.A)
        self.solid = np.zeros(self.shape, np.uint8) if solid is None else \
            np.ascontiguousarray(np.asarray(solid, np.uint8).reshape(self.shape))
        ct = np.float64 if self.cfg["storage"] == 3 else np.float32
        self.macros_arr = np.zeros((4,) + self.shape, ct)

    def init_equilibrium(self, rho, ux, uy, uz=0.0):
        # fields are first rounded to float32, as the GPU receives them
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float32), self.shape).astype(np.float64))
                for a in (rho, ux, uy, uz)]
        lib.orc_pull_init_eq(C.byref(self.c), *[P(a) for a in arrs], P(self.A))
        self.B[...] = self.A
        self.c.step = 0

    def set_populations(self, raw):
        self.A[...] = np.asarray(raw).reshape(self.A.shape)
        self.B[...] = self.A
        self.c.step = 0

    def run(self, steps):
        lib.orc_pull_run(C.byref(self.c), P(self.A), P(self.B), P(self.solid), int(steps), P(self.macros_arr))

    def macros(self):
        return dict(zip(("rho", "ux", "uy", "uz"), self.macros_arr))

    def decoded(self):
        out = np.empty(self.A.shape, np.float64)
        lib.orc_pull_decode(C.byref(self.c), P(self.A), P(out))
        return out

    def reduce(self):
        out = np.zeros(4, np.float64)
        lib.orc_pull_reduce(C.byref(self.c), P(self.A), P(self.solid), P(out))
        return out


# ---------------------------------------------------------------------------
# Reference transcription: the runLBM driver over ndev subdomains
# ---------------------------------------------------------------------------
class RefOracle:
    def __init__(self, solids, deviceCount=1, ux=0.0, uy=0.0, uz=0.0, rho=1.0, viscosity=1.0 / 6.0,
                 Fx=0.0, Fy=0.0, G=0.0, Gads=0.0, cursor=(0, 0, 0), zou_he=1):
        s = np.ascontiguousarray(np.asarray(solids, np.uint8))
        if s.ndim == 2:
            s = s[None]
        self.D, self.H, Wf = s.shape
        self.nd = int(deviceCount)
        self.W = Wf // self.nd
        self.is3d = self.D > 1
        self.Q = 19 if self.is3d else 9
        N = self.W * self.H * self.D
        self.N = N
        nd = self.nd
        self.f = np.zeros((nd, self.Q, N), np.float32)
        self.ux, self.uy, self.uz, self.rho = (np.zeros((nd, N), np.float32) for _ in range(4))
        self.psi, self.ueq_x, self.ueq_y = (np.zeros((nd, N), np.float32) for _ in range(3))
        self.solid = np.zeros((nd, N), np.uint8)
        self.sb = np.zeros((nd, N), np.uint8)
        self.plot = np.zeros((nd, (4 if self.is3d else 8) * N), np.float32)
        self.tau_inv = np.float32(1.0 / (3.0 * np.float64(np.float32(viscosity)) + 0.5))
        self.p = dict(ux=ux, uy=uy, uz=uz, rho=rho, Fx=Fx, Fy=Fy, G=G, Gads=Gads)
        self.cursor = cursor
        self.zou_he = zou_he
        self.value_cur = np.zeros(4, np.float32)
        if self.is3d:
            rc = lib.orc_ref_init_d3q19(P(s), Wf, self.H, self.D, nd, ux, uy, uz, rho, P(self.f), P(self.ux),
                                        P(self.uy), P(self.uz), P(self.rho), P(self.solid), P(self.sb))
        else:
            rc = lib.orc_ref_init_d2q9(P(s), Wf, self.H, nd, ux, uy, rho, float(self.tau_inv), Fx, Fy, G, Gads,
                                       P(self.f), P(self.ux), P(self.uy), P(self.rho), P(self.psi), P(self.ueq_x),
                                       P(self.ueq_y), P(self.solid), P(self.sb))
        if rc:
            raise ValueError("reference initialiser does not support this deck")

    def runLBM(self, steps):
        p = self.p
        cx, cy, cz = self.cursor
        vc = self.value_cur
        if self.is3d:
            lib.orc_ref_run_d3q19(self.nd, self.W, self.H, self.D, P(self.f), P(self.solid), P(self.sb), P(self.ux),
                                  P(self.uy), P(self.uz), P(self.rho), P(self.plot), p["ux"], p["uy"], p["uz"],
                                  p["rho"], float(self.tau_inv), int(steps), cx, cy, cz, P(vc), self.zou_he)
        else:
            lib.orc_ref_run_d2q9(self.nd, self.W, self.H, P(self.f), P(self.solid), P(self.sb), P(self.ux),
                                 P(self.uy), P(self.rho), P(self.psi), P(self.ueq_x), P(self.ueq_y), P(self.plot),
                                 p["ux"], p["uy"], p["rho"], float(self.tau_inv), p["Fx"], p["Fy"], p["G"],
                                 p["Gads"], int(steps), cx, cy, P(vc), self.zou_he)
