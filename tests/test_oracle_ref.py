"""Pins of the reference transcription (orc_ref_*) by properties that a
plausible transcription mistake breaks: streaming is the permutation x -> x+c_i,
BGK conserves per-node mass and momentum, Zou-He reproduces the prescribed
boundary values, k subdomains with halo exchange equal one domain, and the
initial equilibrium has the prescribed moments."""
import numpy as np

import oracle as O

C9 = np.array([[0, 0], [1, 0], [0, -1], [-1, 0], [0, 1], [1, -1], [-1, -1], [-1, 1], [1, 1]])
C19 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                [1, -1, 0], [1, 0, 1], [1, 0, -1], [-1, 1, 0], [-1, -1, 0], [-1, 0, 1], [-1, 0, -1],
                [0, 1, 1], [0, 1, -1], [0, -1, 1], [0, -1, -1]])
OPP9 = [0, 3, 4, 1, 2, 7, 8, 5, 6]
OPP19 = [0, 2, 1, 4, 3, 6, 5, 12, 11, 14, 13, 8, 7, 10, 9, 18, 17, 16, 15]


def test_streaming_d2q9_moves_delta_by_c():
    W, H = 8, 6
    N = W * H
    for i in range(1, 9):
        f = np.zeros((9, H, W), np.float32)
        f[i, 2, 3] = 1.0
        sb = np.zeros(N, np.uint8)
        O.lib.orc_ref_streaming_d2q9(O.P(f), O.P(sb), W, H)
        # after the swap the arrived population sits in the opposite slot
        j = OPP9[i]
        x, y = (3 + C9[i][0]) % W, (2 + C9[i][1]) % H
        assert f[j, y, x] == 1.0 and f.sum() == 1.0, i


def test_streaming_d3q19_moves_delta_by_c_and_preserves_sums():
    W, H, D = 5, 4, 3
    N = W * H * D
    for i in range(1, 19):
        f = np.zeros((19, D, H, W), np.float32)
        f[i, 1, 2, 3] = 1.0
        O.lib.orc_ref_streaming_d3q19(O.P(f), O.P(np.zeros(N, np.uint8)), W, H, D)
        j = OPP19[i]
        x, y, z = (3 + C19[i][0]) % W, (2 + C19[i][1]) % H, (1 + C19[i][2]) % D
        assert f[j, z, y, x] == 1.0 and f.sum() == 1.0, i
    rng = np.random.default_rng(0)
    f = rng.uniform(0, 1, (19, N)).astype(np.float32)
    before = np.sort(f.ravel())
    O.lib.orc_ref_streaming_d3q19(O.P(f), O.P(np.zeros(N, np.uint8)), W, H, D)
    np.testing.assert_array_equal(np.sort(f.ravel()), before)


def _post_stream_like(rng, Q, N, w):
    # populations near equilibrium at rest with a small random perturbation
    return (w[:, None] * (1.0 + 0.05 * rng.standard_normal((Q, N)))).astype(np.float32)


def test_collision_d3q19_conserves_mass_momentum_per_node():
    rng = np.random.default_rng(1)
    W, H, D = 6, 4, 3
    N = W * H * D
    f = _post_stream_like(rng, 19, N, O.W19F.astype(np.float64))
    cphys = C19[np.array(OPP19)]       # slots are swapped before the moments
    m0 = f.astype(np.float64).sum(0)
    p0 = (f.astype(np.float64)[:, None, :] * cphys[:, :, None]).sum(0)
    z = np.zeros(N, np.float32)
    plot = np.zeros(4 * N, np.float32)
    ux, uy, uz, rho = z.copy(), z.copy(), z.copy(), z.copy()
    O.lib.orc_ref_collision_d3q19(O.P(plot), O.P(f), O.P(np.zeros(N, np.uint8)), O.P(ux), O.P(uy), O.P(uz),
                                  O.P(rho), W, H, D, 0, 1, *O.W19F[[0, 1, 7]].tolist(), 1 / 0.56, 0.0, 0.0, 0.0,
                                  1.0, 0)
    m1 = f.astype(np.float64).sum(0)
    p1 = (f.astype(np.float64)[:, None, :] * C19[:, :, None]).sum(0)
    np.testing.assert_allclose(m1, m0, rtol=2e-6)
    np.testing.assert_allclose(p1, p0, atol=2e-7)
    np.testing.assert_allclose(rho, m0, rtol=1e-6)
    np.testing.assert_array_equal(plot.reshape(N, 4)[:, 3], rho)


def test_zou_he_reproduces_inlet_velocity_and_outlet_density_d2q9():
    rng = np.random.default_rng(2)
    W, H = 8, 4
    N = W * H
    f = _post_stream_like(rng, 9, N, O.W9F.astype(np.float64))
    z = np.zeros(N, np.float32)
    ux, uy, rho, psi, ex, ey = (z.copy() for _ in range(6))
    plot = np.zeros(8 * N, np.float32)
    u_in, rho_out = 0.04, 1.02
    O.lib.orc_ref_collision_d2q9(O.P(plot), O.P(f), O.P(np.zeros(N, np.uint8)), O.P(ux), O.P(uy), O.P(rho),
                                 O.P(psi), O.P(ex), O.P(ey), W, H, 0, 1, *O.W9F[[0, 1, 5]].tolist(), 1.2,
                                 u_in, 0.0, rho_out, 0.0, 0.0, 0.0, 0.0, 1)
    ux, uy, rho = ux.reshape(H, W), uy.reshape(H, W), rho.reshape(H, W)
    np.testing.assert_allclose(ux[:, 0], u_in, atol=2e-7)
    np.testing.assert_allclose(uy[:, 0], 0.0, atol=2e-7)      # transverse closure
    np.testing.assert_allclose(rho[:, -1], rho_out, atol=3e-7)
    np.testing.assert_allclose(uy[:, -1], 0.0, atol=2e-7)


def test_zou_he_d3q19_inlet_velocity_outlet_density():
    rng = np.random.default_rng(3)
    W, H, D = 6, 3, 3
    N = W * H * D
    f = _post_stream_like(rng, 19, N, O.W19F.astype(np.float64))
    z = np.zeros(N, np.float32)
    ux, uy, uz, rho = (z.copy() for _ in range(4))
    plot = np.zeros(4 * N, np.float32)
    O.lib.orc_ref_collision_d3q19(O.P(plot), O.P(f), O.P(np.zeros(N, np.uint8)), O.P(ux), O.P(uy), O.P(uz),
                                  O.P(rho), W, H, D, 0, 1, *O.W19F[[0, 1, 7]].tolist(), 1.5, 0.03, 0.0, 0.0,
                                  0.98, 1)
    ux, uy, rho = ux.reshape(D, H, W), uy.reshape(D, H, W), rho.reshape(D, H, W)
    np.testing.assert_allclose(ux[:, :, 0], 0.03, atol=3e-7)
    np.testing.assert_allclose(uy[:, :, 0], 0.0, atol=3e-7)
    np.testing.assert_allclose(rho[:, :, -1], 0.98, atol=3e-7)


def test_init_moments_d2q9_and_d3q19():
    s2 = np.zeros((4, 8), np.uint8)
    r = O.RefOracle(s2, ux=0.05, uy=-0.02, rho=1.1)
    f = r.f[0].astype(np.float64)
    np.testing.assert_allclose(f.sum(0), 1.1, rtol=1e-6)
    cphys = C9 * np.array([1, -1])
    np.testing.assert_allclose((f * cphys[:, 0:1]).sum(0), 1.1 * 0.05, rtol=1e-5)
    np.testing.assert_allclose((f * cphys[:, 1:2]).sum(0), 1.1 * -0.02, rtol=1e-5)
    s3 = np.zeros((2, 3, 4), np.uint8)
    r3 = O.RefOracle(s3, ux=0.03, uy=0.01, uz=-0.02, rho=0.9)
    f3 = r3.f[0].astype(np.float64)
    np.testing.assert_allclose(f3.sum(0), 0.9, rtol=1e-6)
    for a, u in enumerate((0.03, 0.01, -0.02)):
        np.testing.assert_allclose((f3 * C19[:, a:a + 1]).sum(0), 0.9 * u, rtol=1e-5)


def _solids2d(W, H):
    s = np.zeros((H, W), np.uint8)
    s[0, :] = 1
    s[-1, :] = 1
    s[5:8, 10:13] = 1
    return s


def test_halo_exchange_subdomains_equal_single_domain_d2q9():
    W, H = 32, 12
    s = _solids2d(W, H)
    one = O.RefOracle(s, deviceCount=1, ux=0.05, rho=1.0, viscosity=0.05, zou_he=0)
    four = O.RefOracle(s, deviceCount=4, ux=0.05, rho=1.0, viscosity=0.05, zou_he=0)
    one.runLBM(40)
    four.runLBM(40)
    full = one.f[0].reshape(9, H, W)
    parts = np.concatenate([four.f[d].reshape(9, H, W // 4) for d in range(4)], axis=2)
    fluid = s == 0
    np.testing.assert_array_equal(parts[:, fluid], full[:, fluid])


def test_halo_exchange_subdomains_equal_single_domain_d3q19():
    W, H, D = 16, 6, 4
    s = np.zeros((D, H, W), np.uint8)
    s[1:3, 2:4, 6:9] = 1
    one = O.RefOracle(s, deviceCount=1, ux=0.03, viscosity=0.04, zou_he=0)
    two = O.RefOracle(s, deviceCount=2, ux=0.03, viscosity=0.04, zou_he=0)
    one.runLBM(25)
    two.runLBM(25)
    full = one.f[0].reshape(19, D, H, W)
    parts = np.concatenate([two.f[d].reshape(19, D, H, W // 2) for d in range(2)], axis=3)
    fluid = s == 0
    np.testing.assert_array_equal(parts[:, fluid], full[:, fluid])


def test_border_pack_unpack_roundtrip():
    rng = np.random.default_rng(4)
    W, H, D = 5, 4, 3
    f = rng.standard_normal((19, D * H * W)).astype(np.float32)
    b = np.zeros(10 * H * D, np.float32)
    O.lib.orc_ref_get_border_d3q19(O.P(f), O.P(b), W, H, D)
    g = np.zeros_like(f)
    O.lib.orc_ref_apply_border_d3q19(O.P(g), O.P(b), W, H, D)
    F, G = f.reshape(19, D, H, W), g.reshape(19, D, H, W)
    for i in (1, 7, 8, 9, 10):
        np.testing.assert_array_equal(G[i, :, :, -1], F[i, :, :, -1])
    for i in (2, 11, 12, 13, 14):
        np.testing.assert_array_equal(G[i, :, :, 0], F[i, :, :, 0])
