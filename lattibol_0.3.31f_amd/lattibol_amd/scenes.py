"""Seeded synthetic scenes for the five benchmark workloads.

Randomness comes from SplitMix64 (Steele, Lea & Flood 2014): output k of
stream `seed` is mix(seed + (k+1) * 0x9e3779b97f4a7c15), so any element can be
generated independently and the host, the GPU and the oracle agree exactly.
"""
import numpy as np

from .solver import make_config

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, start=0):
    """Outputs start..start+n-1 of the SplitMix64 stream `seed` (uint64 array)."""
    k = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n, start=0):
    """Doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, n, start) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_noise(seed, n, base, amp):
    """float32(base + amp * (2u - 1)), u = uniform01 -- the noisy density start."""
    return (base + amp * (2.0 * uniform01(seed, n) - 1.0)).astype(np.float32)


# ---------------------------------------------------------------------------
# C1: lid-driven cavity, D2Q9 256x256, tau 0.6, lid u = 0.1 on the top row.
# Storage rows grow downward (as in the reference), so the top row is row 0,
# i.e. face -y.  Half-way bounce-back on the other three walls.
# ---------------------------------------------------------------------------
def cavity(n=256, tau=0.6, lid=0.1, storage="f32", **kw):
    cfg = make_config(9, (n, n), tau=tau, storage=storage,
                      bc={"-x": "wall", "+x": "wall", "-y": "moving_wall", "+y": "wall"},
                      bc_u={"-y": (lid, 0.0, 0.0)}, **kw)
    solid = np.zeros((1, n, n), np.uint8)
    init = dict(rho=1.0, ux=0.0, uy=0.0, uz=0.0)
    return cfg, solid, init


# ---------------------------------------------------------------------------
# C2: D3Q19 channel past a sphere, 256x128x128, r = 16 at (64,64,64),
# equilibrium inlet u = 0.05 at x = 0, zero-gradient outflow at x = 255,
# periodic y, z periodic through the z-slab halo planes, tau 0.56.
# ---------------------------------------------------------------------------
def sphere_channel(nx=256, ny=128, nz=128, radius=16, center=(64, 64, 64), u_in=0.05, tau=0.56,
                   storage="f32", halo_z=True, **kw):
    cfg = make_config(19, (nx, ny, nz), tau=tau, storage=storage,
                      bc={"-x": "inlet", "+x": "outflow", "-y": "periodic", "+y": "periodic",
                          "-z": "periodic", "+z": "periodic"},
                      bc_u={"-x": (u_in, 0.0, 0.0)}, bc_rho={"-x": 1.0}, halo_z=halo_z, **kw)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cx, cy, cz = center
    solid = ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 <= radius * radius).astype(np.uint8)
    init = dict(rho=1.0, ux=u_in, uy=0.0, uz=0.0)
    return cfg, solid, init


# ---------------------------------------------------------------------------
# C3: D3Q19 Shan-Chen spinodal decomposition, 192^3 periodic,
# psi = rho0 (1 - exp(-rho/rho0)), rho0 = 1, G = -5, tau = 1,
# rho = 0.7 + U(-0.01, 0.01) from SplitMix64 stream `seed`.
# ---------------------------------------------------------------------------
def spinodal(n=192, G=-5.0, tau=1.0, rho0=1.0, base=0.7, amp=0.01, seed=0, storage="f32", **kw):
    cfg = make_config(19, (n, n, n), tau=tau, storage=storage, sc_G=G, sc_rho0=rho0, **kw)
    solid = np.zeros((n, n, n), np.uint8)
    rho = uniform_noise(seed, n * n * n, base, amp).reshape(n, n, n)
    init = dict(rho=rho, ux=0.0, uy=0.0, uz=0.0)
    return cfg, solid, init


# ---------------------------------------------------------------------------
# C4: ensemble of D2Q9 channels with circular obstacles.
# ---------------------------------------------------------------------------
SIZES = ((64, 64), (128, 64), (128, 128), (256, 128))   # (nx, ny)


def channel_geometry(nx, ny, circles):
    y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    solid = np.zeros((ny, nx), np.uint8)
    for (cx, cy, r) in circles:
        solid |= ((x - cx) ** 2 + (y - cy) ** 2 <= r * r).astype(np.uint8)
    return solid


def _place_circle(nx, ny, r, u_x, u_y):
    xlo, xhi = r + 4, max(r + 4, nx - r - 8)
    ylo, yhi = r + 2, max(r + 2, ny - r - 3)
    return (xlo + int(u_x * (xhi - xlo + 1)), ylo + int(u_y * (yhi - ylo + 1)), r)


def channel_spec(seed, index, size=None, min_open=0.5):
    """Lattice `index` of ensemble `seed`: size, 1-4 circles (r 4..16),
    u_in in [0.02, 0.08], tau in [0.52, 0.9]; 16 uniforms per lattice.

    A circle whose addition would leave any column with less than `min_open`
    of its interior rows open is shrunk (r - 1, same uniforms for its centre)
    until it fits; one that does not fit at r = 4 is dropped (the first
    circle is always kept at r = 4 or more).  Without this, overlapping
    circles can close the channel and the velocity inlet then drives the gap
    flow past Ma = 1."""
    u = uniform01(seed, 16, start=16 * index)
    nx, ny = size if size is not None else SIZES[int(u[0] * 4)]
    ncirc = 1 + int(u[1] * 4)
    circles = []
    interior = ny - 2
    for c in range(ncirc):
        r = min(4 + int(u[2 + 3 * c] * 13), (ny - 6) // 2)
        while True:
            cand = circles + [_place_circle(nx, ny, r, u[3 + 3 * c], u[4 + 3 * c])]
            blocked = channel_geometry(nx, ny, cand)[1:-1].sum(axis=0)
            if (interior - blocked).min() >= min_open * interior:
                circles = cand
                break
            if r <= 4:
                if not circles:
                    circles = cand
                break
            r -= 1
    u_in = 0.02 + 0.06 * u[14]
    tau = 0.52 + 0.38 * u[15]
    return dict(nx=nx, ny=ny, tau=float(tau), u_in=float(np.float32(u_in)),
                solid=channel_geometry(nx, ny, circles), circles=circles)


def ensemble(n=512, seed=1234, size=None, start=0):
    return [channel_spec(seed, start + i, size) for i in range(n)]


def channel_config(spec, storage="f32s", **kw):
    """Single-lattice configuration equal to one member of the batch: the
    reference's Zou-He velocity inlet (u_in, 0) at x = 0 and pressure outlet
    rho = 1 at x = nx-1, half-way bounce-back walls on rows 0 and ny-1, FP32
    deviation storage (s_i = f_i - w_i, collision in deviation form)."""
    cfg = make_config(9, (spec["nx"], spec["ny"]), tau=spec["tau"], storage=storage,
                      bc={"-x": "outflow", "+x": "outflow", "-y": "wall", "+y": "wall"},
                      zou_he=True, zh_ux_in=spec["u_in"], zh_rho_out=1.0, rest_closure=True, **kw)
    return cfg
