"""GPU parity: liblattibol_amd.so on the MI355X against the CPU oracle.

Bar: integer/byte work bit-exact; FP32/FP64 populations bit-identical to the
oracle (the kernels evaluate the same IEEE operations in the same order with
no FMA contraction), and at least within 1e-5 relative after 1000 FP32 steps
as required; FP16/BF16 stored values identical to the oracle's emulation.
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

O.set_threads(min(16, os.cpu_count() or 1))


def _max_rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))


def _assert_pop_parity(gpu_raw, orc_raw, storage, fluid=None, rel=1e-5):
    g, o = gpu_raw, orc_raw
    if fluid is not None:
        g, o = g[:, fluid], o[:, fluid]
    if storage in (1, 2):                      # 16-bit stored deviations
        np.testing.assert_array_equal(g, o)
        return
    nbad = int(np.count_nonzero(g != o))
    assert _max_rel(g, o) <= rel, (nbad, _max_rel(g, o))
    assert nbad == 0, "populations within %.1e but %d values differ bitwise" % (_max_rel(g, o), nbad)


# ---------------------------------------------------------------------------
# Reference driver (streaming / halo / collision / probe launchers)
# ---------------------------------------------------------------------------
REF_DECKS = {
    "d2q9_zouhe": dict(shape=(32, 64), ndev=1, kw=dict(ux=0.05, rho=1.0, viscosity=0.05), steps=200),
    "d2q9_4dev": dict(shape=(32, 128), ndev=4, kw=dict(ux=0.04, rho=1.0, viscosity=0.08), steps=120),
    "d2q9_force": dict(shape=(16, 64), ndev=1, kw=dict(ux=0.0, rho=1.0, viscosity=0.1, Fx=2.035e-6), steps=150),
    "d2q9_shanchen": dict(shape=(32, 64), ndev=1, kw=dict(ux=0.0, rho=200.0, viscosity=1 / 6, G=-120.0,
                                                         Gads=-100.0), steps=150),
    "d3q19_zouhe": dict(shape=(8, 12, 32), ndev=1, kw=dict(ux=0.04, rho=1.0, viscosity=0.04), steps=80),
    "d3q19_2dev": dict(shape=(8, 12, 32), ndev=2, kw=dict(ux=0.03, rho=1.0, viscosity=0.05), steps=60),
}


def _deck_solids(shape):
    s = np.zeros(shape, np.uint8)
    if len(shape) == 2:
        H, W = shape
        s[0, :] = s[-1, :] = 1
        s[H // 3:H // 3 + 4, W // 4:W // 4 + 5] = 1
    else:
        D, H, W = shape
        s[:, 0, :] = s[:, -1, :] = 1
        s[2:5, 4:7, 6:10] = 1
    return s


@pytest.mark.parametrize("deck", sorted(REF_DECKS))
def test_reference_driver_matches_oracle(gpu, deck):
    d = REF_DECKS[deck]
    s = _deck_solids(d["shape"])
    cursor = (3, 5, 1 if len(d["shape"]) == 3 else 0)
    sim = gpu.Simulation(s, deviceCount=d["ndev"], cursor=cursor, **d["kw"])
    ref = O.RefOracle(s, deviceCount=d["ndev"], cursor=cursor, **d["kw"])
    sim.runLBM(d["steps"])
    ref.runLBM(d["steps"])
    for dev in range(d["ndev"]):
        got = sim.download(dev)
        np.testing.assert_array_equal(got["f"].reshape(ref.Q, -1), ref.f[dev])
        np.testing.assert_array_equal(got["rho"].ravel(), ref.rho[dev])
        np.testing.assert_array_equal(got["ux"].ravel(), ref.ux[dev])
        np.testing.assert_array_equal(got["plot"], ref.plot[dev])
    np.testing.assert_array_equal(sim.valueCur(), ref.value_cur)
    sim.close()


def test_reference_launchers_direct_and_shape_check(gpu):
    import ctypes as C
    from lattibol_amd import _lib
    W, H = 64, 16
    s = _deck_solids((H, W))
    sim = gpu.Simulation(s, ux=0.05, viscosity=0.05)
    sim.runLBM(3)
    before = sim.download()["f"].reshape(9, -1).copy()
    data = sim.data()
    grid, block = _lib.LbDim3(W // 32, H // 4, 1), _lib.LbDim3(32, 4, 1)
    _lib.lib.lb_last_error()
    _lib.lib.streaming_d2q9(C.byref(data), grid, block)
    _lib.check_sticky("streaming_d2q9")
    after = sim.download()["f"].reshape(9, -1)
    sb = np.zeros(W * H, np.uint8)
    O.lib.orc_ref_solid_bound_d2q9(O.P(np.ascontiguousarray(s)), O.P(sb), W, H)
    O.lib.orc_ref_streaming_d2q9(O.P(before), O.P(sb), W, H)
    np.testing.assert_array_equal(after, before)
    # grid*block that does not describe the lattice: rejected, nothing launched
    _lib.lib.streaming_d2q9(C.byref(data), _lib.LbDim3(1, 1, 1), block)
    assert _lib.lib.lb_last_error() == -2
    np.testing.assert_array_equal(sim.download()["f"].reshape(9, -1), after)
    # read_value_2d synchronises and fills valueCur from the plot buffer
    plot = sim.download()["plot"]
    _lib.lib.read_value_2d(data.plotBuffer, C.byref(data), 7, 9, 0)
    _lib.check_sticky("read_value_2d")
    np.testing.assert_array_equal(sim.valueCur(), plot.reshape(-1, 4)[7 + 9 * W])
    sim.close()


# ---------------------------------------------------------------------------
# Fused pull-scheme solver
# ---------------------------------------------------------------------------
def _run_pair(gpu, cfg, solid, init, steps):
    sol = gpu.Solver(cfg, solid)
    orc = O.PullOracle(cfg, solid)
    sol.init_equilibrium(init["rho"], init["ux"], init["uy"], init.get("uz", 0.0))
    orc.init_equilibrium(init["rho"], init["ux"], init["uy"], init.get("uz", 0.0))
    np.testing.assert_array_equal(sol.populations(), orc.A)
    sol.step(steps)
    orc.run(steps)
    return sol, orc


FUSED_CASES = {
    "cavity64_f32_1000": ("cavity", dict(n=64), 1000),
    "cavity48_f16": ("cavity", dict(n=48, storage="f16"), 300),
    "cavity48_bf16": ("cavity", dict(n=48, storage="bf16"), 300),
    "cavity32_f64": ("cavity", dict(n=32, storage="f64"), 300),
    "sphere_halo_f32_1000": ("sphere", dict(nx=64, ny=32, nz=32, radius=5, center=(16, 16, 16)), 1000),
    "sphere_nohalo_f32": ("sphere", dict(nx=48, ny=24, nz=16, radius=4, center=(12, 12, 8), halo_z=False), 200),
    "sphere_halo_bf16": ("sphere", dict(nx=48, ny=24, nz=16, radius=4, center=(12, 12, 8), storage="bf16"), 200),
    "sphere_halo_f16": ("sphere", dict(nx=48, ny=24, nz=16, radius=4, center=(12, 12, 8), storage="f16"), 200),
    "spinodal32_f32_1000": ("spinodal", dict(n=32), 1000),
    "spinodal24_f16": ("spinodal", dict(n=24, storage="f16"), 300),
    "spinodal24_bf16": ("spinodal", dict(n=24, storage="bf16"), 300),
    "spinodal24_f16_nearest": ("spinodal", dict(n=24, storage="f16", round_mode="nearest"), 300),
    "spinodal20_f64": ("spinodal", dict(n=20, storage="f64"), 300),
    "cavity48_f32s": ("cavity", dict(n=48, storage="f32s"), 500),
    "sphere_halo_f32s": ("sphere", dict(nx=48, ny=24, nz=16, radius=4, center=(12, 12, 8), storage="f32s"), 300),
    "spinodal24_f32s": ("spinodal", dict(n=24, storage="f32s"), 300),
    "channel_zouhe_f32s": ("channel", dict(seed=3, index=0, size=(128, 64)), 1000),
    "cavity48_bf16_sr": ("cavity", dict(n=48, storage="bf16", round_mode=1), 300),
    "sphere_halo_f16_sr": ("sphere", dict(nx=48, ny=24, nz=16, radius=4, center=(12, 12, 8), storage="f16",
                                          round_mode=1), 200),
    "spinodal24_bf16_sr": ("spinodal", dict(n=24, storage="bf16", round_mode=1), 300),
    # Shan-Chen beside walls, a moving wall, an outflow face and solids
    # (psi clamped at non-periodic faces, sc_psi_solid on solid nodes), ragged
    # tile edges; D2Q9 and D3Q19
    "sc9_walls_obstacle_f32": ("sc", dict(name="sc9"), 400),
    "sc19_walls_box_f32": ("sc", dict(name="sc19"), 300),
    "sc19_walls_box_bf16_sr": ("sc", dict(name="sc19", storage="bf16", round_mode=1), 200),
}


def _sc_scene(name, storage="f32", round_mode="default"):
    from lattibol_amd import scenes
    from lattibol_amd.solver import make_config
    if name == "sc9":
        nx, ny = 50, 37
        cfg = make_config(9, (nx, ny), tau=1.0, sc_G=-5.0, sc_rho0=1.0, sc_psi_solid=0.4, storage=storage,
                          round_mode=round_mode, bc={"-y": "wall", "+y": "wall"})
        y, x = np.mgrid[0:ny, 0:nx]
        solid = (((x - 20) ** 2 + (y - 15) ** 2) <= 36).astype(np.uint8)[None]
        rho = scenes.uniform_noise(5, nx * ny, 0.7, 0.05).reshape(1, ny, nx)
    else:
        nx, ny, nz = 45, 20, 16
        cfg = make_config(19, (nx, ny, nz), tau=1.0, sc_G=-5.0, sc_rho0=1.0, sc_psi_solid=0.2, storage=storage,
                          round_mode=round_mode,
                          bc={"-y": "wall", "+y": "outflow", "-z": "wall", "+z": "moving_wall"},
                          bc_u={"+z": (0.02, 0.0, 0.0)})
        solid = np.zeros((nz, ny, nx), np.uint8)
        solid[4:9, 5:10, 10:16] = 1
        rho = scenes.uniform_noise(6, nx * ny * nz, 0.7, 0.05).reshape(nz, ny, nx)
    return cfg, solid, dict(rho=rho, ux=0.0, uy=0.0, uz=0.0)


def _channel_scene(seed, index, size):
    from lattibol_amd import scenes
    spec = scenes.channel_spec(seed, index, size)
    return scenes.channel_config(spec), spec["solid"][None], dict(rho=1.0, ux=0.0, uy=0.0, uz=0.0)


def _scene(kind, kw):
    from lattibol_amd import scenes
    return {"cavity": scenes.cavity, "sphere": scenes.sphere_channel, "spinodal": scenes.spinodal,
            "channel": _channel_scene, "sc": _sc_scene}[kind](**kw)


@pytest.mark.parametrize("words", [0, 1])
@pytest.mark.parametrize("layout", ["soa", "rows"])
@pytest.mark.parametrize("case", sorted(FUSED_CASES))
def test_fused_solver_matches_oracle(gpu, case, layout, words):
    """Both device population layouts and both node-flag forms (byte + link
    mask, packed word); populations come back in [i][z][y][x]."""
    kind, kw, steps = FUSED_CASES[case]
    cfg, solid, init = _scene(kind, kw)
    cfg["layout"] = {"soa": 0, "rows": 1}[layout]
    cfg["node_words"] = words
    print("case", case, layout, cfg["nx"], cfg["ny"], cfg["nz"], "storage", cfg["storage"], flush=True)
    sol, orc = _run_pair(gpu, cfg, solid, init, steps)
    assert sol.launch_config["layout"] == cfg["layout"] and sol.launch_config["node_words"] == words
    fluid = solid.reshape(orc.shape) == 0
    _assert_pop_parity(sol.populations(), orc.A, cfg["storage"], fluid)
    m, mo = sol.macros(), orc.macros()
    for k in ("rho", "ux", "uy", "uz"):
        np.testing.assert_array_equal(m[k][fluid], mo[k][fluid].astype(np.float32))
    # FP64 totals: GPU tree reduction vs the oracle's long-double sum
    np.testing.assert_allclose(sol.reduce(), orc.reduce(), rtol=1e-10, atol=1e-9)
    sol.close()


def _tb_scene(kind):
    from lattibol_amd import scenes
    from lattibol_amd.solver import make_config
    if kind == "cavity64":
        return scenes.cavity(n=64, time_block=4)
    if kind == "ragged50x37":
        cfg = make_config(9, (50, 37), tau=0.7, bc={"-x": "wall", "+x": "wall", "-y": "moving_wall", "+y": "wall"},
                          bc_u={"-y": (0.08, 0.0, 0.0)}, time_block=4)
        y, x = np.mgrid[0:37, 0:50]
        solid = (((x - 30) ** 2 + (y - 18) ** 2) <= 25).astype(np.uint8)[None]
        return cfg, solid, dict(rho=1.0, ux=0.0, uy=0.0, uz=0.0)
    # periodic box with an obstacle and a uniform initial flow
    cfg = make_config(9, (40, 24), tau=0.6, bc={f: "periodic" for f in ("-x", "+x", "-y", "+y")}, time_block=4)
    y, x = np.mgrid[0:24, 0:40]
    solid = (((x - 12) ** 2 + (y - 11) ** 2) <= 16).astype(np.uint8)[None]
    return cfg, solid, dict(rho=1.0, ux=0.05, uy=0.01, uz=0.0)


@pytest.mark.parametrize("tb", [4, 8])
@pytest.mark.parametrize("words", [0, 1])
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("kind,steps", [("cavity64", 1003), ("ragged50x37", 501), ("periodic40x24", 402)])
def test_time_blocked_steps_match_oracle(gpu, kind, steps, layout, words, tb):
    """Four or eight steps per launch through the LDS tile: the same
    populations as the oracle's single steps, bit for bit (walls, moving lid,
    obstacles, periodic faces, ragged tiles, step counts not divisible by the
    block)."""
    cfg, solid, init = _tb_scene(kind)
    cfg["layout"], cfg["node_words"], cfg["time_block"] = layout, words, tb
    print("case", kind, layout, words, tb, flush=True)
    sol, orc = _run_pair(gpu, cfg, solid, init, steps)
    assert sol.launch_config["time_block"] == tb
    fluid = solid.reshape(orc.shape) == 0
    _assert_pop_parity(sol.populations(), orc.A, cfg["storage"], fluid)
    m, mo = sol.macros(), orc.macros()
    for k in ("rho", "ux", "uy"):
        np.testing.assert_array_equal(m[k][fluid], mo[k][fluid].astype(np.float32))
    sol.close()


def test_time_block_rejected_where_it_cannot_apply(gpu):
    from lattibol_amd import scenes
    from lattibol_amd._lib import LatticeError
    cfg, solid, init = scenes.sphere_channel(32, 16, 8, radius=2, center=(8, 8, 4), time_block=4)
    with pytest.raises(LatticeError):
        gpu.Solver(cfg, solid)
    cfg, solid, init = scenes.cavity(n=32, storage="f16", time_block=4)
    with pytest.raises(LatticeError):
        gpu.Solver(cfg, solid)
    cfg, solid, init = scenes.cavity(n=20, time_block=8)     # smaller than 16 + 8
    with pytest.raises(LatticeError):
        gpu.Solver(cfg, solid)


def test_stochastic_rounding_stream_continues_across_calls(gpu):
    """The device step counter of the rounding stream survives graph chunks,
    eager steps and separate lb_step calls: 100 + 77 steps equal the oracle's
    177 bit for bit."""
    from lattibol_amd import scenes
    cfg, solid, init = scenes.spinodal(n=20, storage="f16", round_mode=1)
    sol = gpu.Solver(cfg, solid)
    orc = O.PullOracle(cfg, solid)
    sol.init_equilibrium(init["rho"], 0.0, 0.0, 0.0)
    orc.init_equilibrium(init["rho"], 0.0, 0.0, 0.0)
    sol.step(100)
    sol.step(77)
    orc.run(177)
    np.testing.assert_array_equal(sol.populations(), orc.A)
    sol.close()


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("Q", [9, 19])
def test_fused_reference_physics_mode_matches_reference_driver(gpu, Q, layout):
    """bb_full + zou_he: the fused sweep reproduces the reference step
    (populations uploaded in [i][z][y][x] order into either device layout)."""
    from lattibol_amd import make_config
    shape = (16, 48) if Q == 9 else (6, 10, 24)
    s = _deck_solids(shape)
    ref = O.RefOracle(s, ux=0.04, rho=1.0, viscosity=0.05)
    dims = tuple(reversed(shape))
    cfg = make_config(Q, dims, tau_inv=float(ref.tau_inv), bb_full=True, zou_he=True, zh_ux_in=0.04,
                      zh_rho_out=1.0, layout=layout)
    sol = gpu.Solver(cfg, s)
    sol.set_populations(ref.f[0])
    np.testing.assert_array_equal(sol.populations().reshape(ref.f[0].shape), ref.f[0])
    ref.runLBM(150)
    sol.step(150)
    fluid = s.reshape((1,) + shape if Q == 9 else shape) == 0
    got = sol.populations()
    np.testing.assert_array_equal(got[:, fluid], ref.f[0].reshape(got.shape)[:, fluid])
    sol.close()


@pytest.mark.parametrize("layout", [0, 1])
def test_halo_planes_equal_periodic_z(gpu, layout):
    from lattibol_amd import scenes
    outs = []
    for halo in (True, False):
        cfg, solid, init = scenes.sphere_channel(64, 32, 24, radius=6, center=(20, 16, 12), halo_z=halo,
                                                 layout=layout)
        sol = gpu.Solver(cfg, solid)
        sol.init_equilibrium(init["rho"], init["ux"], init["uy"], init["uz"])
        sol.step(150)
        outs.append(sol.populations())
        sol.close()
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("layout", [0, 1])
def test_two_slab_ring_equals_one_lattice(gpu, layout):
    """The z-slab ring protocol on one GPU: two solvers own z = 0..11 and
    12..23 of a periodic-z channel; before every single step the planes each
    one packed (left in its recv regions, see lb_halo_buffers) are swapped
    between them (recv_hi <-> recv_hi, recv_lo <-> recv_lo), which hands each
    slab its neighbour's outgoing planes.  The two slabs must reproduce the
    24-plane lattice bit for bit.  The obstacle stays clear of the slab
    boundaries: each slab's link masks see only its own solids."""
    import ctypes as C
    from lattibol_amd import scenes
    from lattibol_amd._lib import check, lib
    nx, ny, nz, half, steps = 64, 32, 24, 12, 60
    kw = dict(radius=4, center=(20, 16, 6), halo_z=True, layout=layout)
    cfg, solid, init = scenes.sphere_channel(nx, ny, nz, **kw)
    assert not solid[[0, half - 1, half, nz - 1]].any()
    full = gpu.Solver(cfg, solid)
    full.init_equilibrium(init["rho"], init["ux"], init["uy"], init["uz"])
    slabs = []
    for k in range(2):
        cfg_k, _, _ = scenes.sphere_channel(nx, ny, half, **kw)
        s_k = gpu.Solver(cfg_k, np.ascontiguousarray(solid[k * half:(k + 1) * half]))
        s_k.init_equilibrium(init["rho"], init["ux"], init["uy"], init["uz"])
        slabs.append(s_k)
    nbytes = 5 * nx * ny * 4
    host = [np.empty(nbytes, np.uint8) for _ in range(2)]
    hp = [h.ctypes.data_as(C.c_void_p) for h in host]
    for t in range(steps):
        for s_k in slabs:
            s_k.sync()
        hb = [s_k.halo_buffers() for s_k in slabs]
        for key in ("recv_hi", "recv_lo"):
            for k in range(2):
                check(lib.lb_memcpy_d2h(hp[k], C.c_void_p(hb[k][key]), nbytes), "d2h")
            for k in range(2):
                check(lib.lb_memcpy_h2d(C.c_void_p(hb[k][key]), hp[1 - k], nbytes), "h2d")
        for s_k in slabs:
            s_k.step(1)
    full.step(steps)
    ref = full.populations()
    for k, s_k in enumerate(slabs):
        np.testing.assert_array_equal(s_k.populations(), ref[:, k * half:(k + 1) * half], err_msg="slab %d" % k)
        s_k.close()
    full.close()


def test_residual_read_value_and_steps(gpu):
    from lattibol_amd import scenes
    cfg, solid, init = scenes.cavity(n=64)
    sol = gpu.Solver(cfg, solid)
    sol.init_equilibrium(**{k: init[k] for k in ("rho", "ux", "uy")})
    sol.step(10)
    assert sol.residual() == 1.0            # first call: no snapshot yet
    m0 = sol.macros()
    sol.step(10)
    r = sol.residual()
    m1 = sol.macros()
    u0 = np.stack([m0["ux"], m0["uy"]]).reshape(2, -1).astype(np.float32)
    u1 = np.stack([m1["ux"], m1["uy"]]).reshape(2, -1).astype(np.float32)
    ref = O.lib.orc_residual(2, u0.shape[1], O.P(np.ascontiguousarray(u1)), O.P(np.ascontiguousarray(u0)), None)
    assert abs(r - ref) <= 1e-12 * ref
    v = sol.read_value(5, 7)
    np.testing.assert_array_equal(v, [m1["ux"][0, 7, 5], m1["uy"][0, 7, 5], 0.0, m1["rho"][0, 7, 5]])
    assert sol.steps_done == 20
    sol.close()


# ---------------------------------------------------------------------------
# Batched channels
# ---------------------------------------------------------------------------
def test_batch_members_equal_single_lattice_runs(gpu):
    from lattibol_amd import scenes
    specs = scenes.ensemble(6, seed=99)
    b = gpu.Batch(specs)
    b.step(120)
    for i, sp in enumerate(specs):
        cfg = scenes.channel_config(sp)
        sol = gpu.Solver(cfg, sp["solid"])
        sol.init_uniform(1.0, 0.0, 0.0, 0.0)
        sol.step(120)
        np.testing.assert_array_equal(b.populations(i), sol.populations()[:, 0])
        np.testing.assert_array_equal(b.macros(i)["ux"], sol.macros()["ux"][0])
        sol.close()
    # and one member against the oracle
    sp = specs[0]
    orc = O.PullOracle(scenes.channel_config(sp), sp["solid"])
    orc.init_equilibrium(1.0, 0.0, 0.0)
    orc.run(120)
    fluid = sp["solid"] == 0
    np.testing.assert_array_equal(b.populations(0)[:, fluid], orc.A[:, 0][:, fluid])
    b.close()


def test_batch_run_to_steady_retires_lattices(gpu):
    from lattibol_amd import scenes
    specs = scenes.ensemble(4, seed=5, size=(64, 64))
    for sp in specs:
        sp["tau"] = 0.9          # viscous: converges quickly
    b = gpu.Batch(specs)
    steps, res = b.run_to_steady(tol=1e-5, k=100, max_steps=40000)
    assert (res < 1e-5).all(), res
    assert (steps < 40000).all()
    b.close()


# ---------------------------------------------------------------------------
# Full-size workloads: parity for a few steps against the oracle, and
# size-independent properties over longer runs.
# ---------------------------------------------------------------------------
def test_full_size_sphere_channel_parity(gpu):
    from lattibol_amd import scenes
    cfg, solid, init = scenes.sphere_channel()
    print("full-size sphere channel 256x128x128", flush=True)
    sol, orc = _run_pair(gpu, cfg, solid, init, 6)
    _assert_pop_parity(sol.populations(), orc.A, 0, solid == 0)
    sol.close()


ONE_PASS_CASES = {
    "spinodal24_f32": ("spinodal", dict(n=24), 300),
    "spinodal24_bf16_sr": ("spinodal", dict(n=24, storage="bf16", round_mode=1), 300),
    "spinodal24_f16_nearest": ("spinodal", dict(n=24, storage="f16", round_mode="nearest"), 300),
    "sc19_walls_box_f32": ("sc", dict(name="sc19"), 300),
    "sc19_walls_box_bf16_sr": ("sc", dict(name="sc19", storage="bf16", round_mode=1), 200),
    "sc19_walls_box_f16": ("sc", dict(name="sc19", storage="f16"), 200),
}


@pytest.mark.parametrize("layout", ["soa", "rows"])
@pytest.mark.parametrize("bx,by,zc,form", [(32, 16, 5, 1), (64, 8, 8, 2), (128, 4, 16, 1), (32, 8, 7, 2),
                                           (64, 4, 3, 1), (64, 16, 6, 1), (128, 8, 4, 2), (32, 32, 9, 1)])
@pytest.mark.parametrize("case", sorted(ONE_PASS_CASES))
def test_one_pass_shan_chen_sweep_matches_oracle(gpu, case, bx, by, zc, form, layout):
    """The z-marching sweep (psi planes in LDS) against the oracle's density
    pass + sweep: tiles of 32 x 16, 64 x 8, 128 x 4 (512 threads), 32 x 8,
    64 x 4 (256 threads), 64 x 16, 128 x 8 and 32 x 32 nodes (1,024), ragged at the +x / +y faces or wider than the
    lattice, z-chunks that do not divide nz, periodic faces and walls /
    outflow / moving wall / solids, both layouts, the populations kept in
    registers (form 1) or pulled again for the sweep (form 2)."""
    kind, kw, steps = ONE_PASS_CASES[case]
    cfg, solid, init = _scene(kind, kw)
    cfg.update(layout={"soa": 0, "rows": 1}[layout], sc_sweep=form, sc_zchunk=zc, block_x=bx, block_y=by)
    print("case", case, "one-pass", bx, by, zc, form, layout, flush=True)
    sol, orc = _run_pair(gpu, cfg, solid, init, steps)
    lc = sol.launch_config
    assert lc["sc_sweep"] == form and lc["sc_zchunk"] == zc and lc["block_x"] == bx and lc["block_y"] == by, lc
    fluid = solid.reshape(orc.shape) == 0
    _assert_pop_parity(sol.populations(), orc.A, cfg["storage"], fluid)
    m, mo = sol.macros(), orc.macros()
    for k in ("rho", "ux", "uy", "uz"):
        np.testing.assert_array_equal(m[k][fluid], mo[k][fluid].astype(np.float32))
    sol.close()


def test_one_pass_sweep_rejected_where_it_does_not_apply(gpu):
    from lattibol_amd import _lib, scenes
    for kw, over in ((dict(n=16), dict(sc_sweep=1, storage=3)),          # FP64 storage
                     (dict(n=16), dict(sc_sweep=1, mem_mode=2)),         # non-temporal accesses
                     (dict(n=16), dict(sc_sweep=1, block_x=256)),        # no tile of 256 x 2
                     (dict(n=16), dict(sc_sweep=1, block_x=32, block_y=64))):   # 2,048 threads
        cfg, solid, init = scenes.spinodal(**kw)
        cfg.update(over)
        with pytest.raises(RuntimeError, match="unsupported|UNSUPPORTED|-3"):
            gpu.Solver(cfg, solid)
    cfg, solid, init = scenes.cavity(n=32)
    cfg.update(sc_sweep=1)
    with pytest.raises(RuntimeError):
        gpu.Solver(cfg, solid)


ROW_SLAB_CASES = {
    # shape (nx, ny, nz), storage, rounding, body force, steps
    "cube24_f32": ((24, 24, 24), "f32", "default", None, 300),
    "cube24_bf16_sr": ((24, 24, 24), "bf16", "default", None, 200),
    "cube24_f16_nearest": ((24, 24, 24), "f16", "nearest", None, 200),
    "ragged70x13x11_f32_force": ((70, 13, 11), "f32", "default", (1e-5, -2e-5, 3e-5), 200),
    "ragged130x5x3_f16_sr": ((130, 5, 3), "f16", "default", None, 150),
    "thin64x7x1_bf16_nearest": ((64, 7, 1), "bf16", "nearest", None, 150),
}


def _periodic_box(shape, storage, round_mode, force):
    from lattibol_amd import scenes
    from lattibol_amd.solver import make_config
    nx, ny, nz = shape
    kw = dict(force=force) if force else {}
    cfg = make_config(19, shape, tau=1.0, sc_G=-5.0, sc_rho0=1.0, storage=storage, round_mode=round_mode, **kw)
    solid = np.zeros((nz, ny, nx), np.uint8)
    rho = scenes.uniform_noise(11, nx * ny * nz, 0.7, 0.05).reshape(nz, ny, nx)
    return cfg, solid, dict(rho=rho, ux=0.0, uy=0.0, uz=0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_shan_chen_tuning_leaves_zeroed_buffers(gpu, storage):
    """lb_create times the Shan-Chen forms on a uniform equilibrium it writes
    into the population buffers; afterwards the buffers are zero again, as
    for every other lattice, and a run from the caller's initial state
    matches the oracle."""
    from lattibol_amd import scenes
    cfg, solid, init = scenes.spinodal(n=24, storage=storage)
    sol = gpu.Solver(cfg, solid)
    assert not np.any(sol.populations())
    sol.close()
    sol, orc = _run_pair(gpu, cfg, solid, init, 40)
    _assert_pop_parity(sol.populations(), orc.A, cfg["storage"])
    sol.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["cavity48_f32", "cavity48_bf16", "sphere_halo_f32"])
def test_solid_mask_changed_mid_run_matches_oracle(gpu, case):
    """lb_set_solid right after an lb_step that returned with its step graphs
    still queued (the call drains the solver's stream before it drops them);
    the run continues on the new mask exactly as the oracle does."""
    from lattibol_amd import scenes
    if case.startswith("cavity"):
        cfg, solid, init = scenes.cavity(n=48, storage=case.split("_")[1])
        solid2 = solid.copy()
        solid2[0, 20:26, 14:30] = 1
    else:
        cfg, solid, init = scenes.sphere_channel(nx=48, ny=24, nz=16, radius=4, center=(12, 12, 8))
        solid2 = solid.copy()
        solid2[6:10, 8:14, 30:36] = 1
    sol, orc = _run_pair(gpu, cfg, solid, init, 130)
    sol.set_solid(solid2)
    orc.solid[...] = solid2.reshape(orc.shape)
    sol.step(70)
    orc.run(70)
    fluid = solid2.reshape(orc.shape) == 0
    _assert_pop_parity(sol.populations(), orc.A, cfg["storage"], fluid)
    sol.close()


ROW_SLAB_SHAPES = [(384, 2, 5), (512, 5, 3), (768, 4, 8), (1024, 5, 16), (512, 1, 1), (768, 2, 3)]
# tiles holding two or three interior segments per wave (BF16 only)
ROW_SLAB_BF16_SHAPES = [("cube24_bf16_sr", 512, 20, 4), ("cube24_bf16_sr", 384, 9, 2),
                        ("thin64x7x1_bf16_nearest", 768, 20, 6), ("thin64x7x1_bf16_nearest", 384, 12, 1)]


@pytest.mark.parametrize("layout", ["soa", "rows"])
@pytest.mark.parametrize("case,threads,by,zc",
                         [(c,) + t for c in sorted(ROW_SLAB_CASES) for t in ROW_SLAB_SHAPES] + ROW_SLAB_BF16_SHAPES)
def test_row_slab_shan_chen_sweep_matches_oracle(gpu, case, threads, by, zc, layout):
    """The row-slab one-pass sweep (whole x-rows per workgroup, the frame rows
    spread over the waves, psi planes in LDS) against the oracle's density
    pass + sweep: rows of 24 to 130 nodes (one to three 64-node segments, the
    last one partial), workgroups of 6 to 16 waves with one to three interior
    segments per wave, tiles of 1 to 20 rows that do not divide ny, z-chunks
    of 1 to 16 planes that do not divide nz, a single z-plane, a body force,
    FP32 and 16-bit storage with both roundings, both layouts, several
    graph-replayed launches."""
    shape, storage, rm, force, steps = ROW_SLAB_CASES[case]
    cfg, solid, init = _periodic_box(shape, storage, rm, force)
    cfg.update(layout={"soa": 0, "rows": 1}[layout], sc_sweep=3, sc_zchunk=zc, block_x=threads, block_y=by)
    print("case", case, "row-slab", threads, by, zc, layout, flush=True)
    sol, orc = _run_pair(gpu, cfg, solid, init, steps)
    lc = sol.launch_config
    assert lc["sc_sweep"] == 3 and lc["sc_zchunk"] == zc and lc["block_x"] == threads and lc["block_y"] == by, lc
    _assert_pop_parity(sol.populations(), orc.A, cfg["storage"])
    m, mo = sol.macros(), orc.macros()
    for k in ("rho", "ux", "uy", "uz"):
        np.testing.assert_array_equal(m[k], mo[k].astype(np.float32))
    sol.close()


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("threads,by,zc", [(512, 5, 3), (1024, 5, 16)])
def test_row_slab_stochastic_rounding_near_unit_density(gpu, storage, threads, by, zc):
    """Density within 2e-5 of 1: the stored deviations are of order 1e-7,
    FP16 subnormals, so the row slab's paired FP16 stochastic stores take
    their integer path in every wave (BF16 keeps its instruction path); both
    stay bit-identical to the oracle."""
    from lattibol_amd import scenes
    shape = (70, 13, 11)
    cfg, solid, init = _periodic_box(shape, storage, "default", None)
    init["rho"] = scenes.uniform_noise(5, int(np.prod(shape)), 1.0, 2e-5).reshape(shape[::-1])
    cfg.update(layout=1, sc_sweep=3, sc_zchunk=zc, block_x=threads, block_y=by)
    print("case near-1 density", storage, threads, by, zc, flush=True)
    sol, orc = _run_pair(gpu, cfg, solid, init, 100)
    assert sol.launch_config["sc_sweep"] == 3 and sol.launch_config["round_mode"] == "stochastic"
    raw = orc.A.view(np.uint16)
    if storage == "f16":
        assert np.mean((raw & 0x7c00) == 0) > 0.5       # mostly subnormal (or zero) stored values
    _assert_pop_parity(sol.populations(), orc.A, cfg["storage"])
    sol.close()


def test_row_slab_sweep_applies_only_to_periodic_boxes(gpu):
    from lattibol_amd import scenes
    # tuned choice on a periodic box: solids set afterwards hand the step to
    # the two passes, which still match the oracle
    cfg, solid, init = _periodic_box((40, 12, 10), "f32", "default", None)
    cfg.update(sc_sweep=-1)
    solid = solid.copy()
    solid[3:5, 4:7, 10:14] = 1
    sol, orc = _run_pair(gpu, cfg, solid, init, 60)
    assert sol.launch_config["sc_sweep"] != 3, sol.launch_config
    fluid = solid.reshape(orc.shape) == 0
    _assert_pop_parity(sol.populations(), orc.A, 0, fluid)
    sol.close()
    # explicit choice: a solid node, a wall face, FP64 storage, a workgroup
    # or a tile the form does not have are refused
    cfg, solid, init = _periodic_box((40, 12, 10), "f32", "default", None)
    cfg.update(sc_sweep=3)
    solid[2, 2, 2] = 1
    with pytest.raises(RuntimeError):
        gpu.Solver(cfg, solid)
    for over in (dict(bc=[0, 0, 1, 1, 0, 0]), dict(storage=3), dict(block_x=256), dict(block_x=512, block_y=40),
                 dict(mem_mode=2)):
        cfg, solid, init = _periodic_box((40, 12, 10), "f32", "default", None)
        cfg.update(sc_sweep=3, **over)
        with pytest.raises(RuntimeError):
            gpu.Solver(cfg, solid)
    cfg, solid, init = scenes.spinodal(n=16)
    cfg.update(sc_sweep=3, block_x=768, block_y=2, sc_zchunk=4)
    sol = gpu.Solver(cfg, solid)
    assert sol.launch_config["sc_sweep"] == 3
    sol.close()


def test_full_size_spinodal_parity_and_mass(gpu):
    from lattibol_amd import scenes
    cfg, solid, init = scenes.spinodal()
    print("full-size spinodal 192^3", flush=True)
    sol, orc = _run_pair(gpu, cfg, solid, init, 3)
    _assert_pop_parity(sol.populations(), orc.A, 0)
    m0 = orc.reduce()[0]
    sol.step(500)
    m1 = sol.reduce()
    assert abs(m1[0] - m0) / m0 < 1e-6, (m1[0], m0)
    sol.close()
