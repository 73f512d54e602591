"""Bit-level pins of the oracle's reference-mode per-node arithmetic
(orc_ref_collision_d2q9 / _d3q19: moments, body-force shift, Zou-He inlet and
outlet, BGK) against a numpy restatement written from the reference kernels'
expressions (collision.cu:193-330 for D2Q9, :375-541 for D3Q19).

numpy evaluates float32 arrays with one IEEE rounding per operation and no
fused multiply-add, and Python float scalars stay weak (they do not widen a
float32 operand), so an expression written in the reference's left-to-right
order with the reference's float / double promotions gives the reference's
bits.  The C++ oracle and the HIP kernels share their per-node expressions;
this restatement is independent of both, so a reordered sum, a dropped
promotion or a swapped population in the shared code fails here."""
import numpy as np

import oracle as O

f32, f64 = np.float32, np.float64


def d(a):
    return a.astype(f64)


def ref_collide_d2q9(f, solid, ti, ux_in, rho_out, Fx, Fy, W):
    """collision.cu:168-330 on every node, vectorised; f is [9][N] float32
    as the streaming kernel leaves it (arrived populations in opposite slots)."""
    f = [f[i].copy() for i in range(9)]
    # "Bounce-Back" swap back to natural slots (collision.cu:193-208)
    f[1], f[3] = f[3].copy(), f[1].copy()
    f[2], f[4] = f[4].copy(), f[2].copy()
    f[5], f[7] = f[7].copy(), f[5].copy()
    f[6], f[8] = f[8].copy(), f[6].copy()
    x = np.arange(f[0].size) % W
    ui, ro = f32(ux_in), f32(rho_out)
    west, east = x == 0, x == W - 1
    # Zou-He velocity inlet, double-promoted (collision.cu:211-221)
    rho_w = ((d(f[0] + f[2] + f[4]) + 2. * d(f[3] + f[7] + f[6])) / (1. - f64(ui))).astype(f32)
    ru = rho_w * ui
    f1w = (d(f[3]) + (2. / 3.) * d(ru)).astype(f32)
    f5w = (d(f[7]) + (1. / 6.) * d(ru) - 0.5 * d(f[2] - f[4])).astype(f32)
    f8w = (d(f[6]) + (1. / 6.) * d(ru) - 0.5 * d(f[4] - f[2])).astype(f32)
    # Zou-He pressure outlet (collision.cu:252-262)
    ux_e = (-1. + (d(f[0] + f[2] + f[4]) + 2. * d(f[1] + f[5] + f[8])) / f64(ro)).astype(f32)
    ru = ro * ux_e
    f3e = (d(f[1]) - (2. / 3.) * d(ru)).astype(f32)
    f7e = (d(f[5]) - (1. / 6.) * d(ru) + 0.5 * d(f[2] - f[4])).astype(f32)
    f6e = (d(f[8]) - (1. / 6.) * d(ru) + 0.5 * d(f[4] - f[2])).astype(f32)
    f[1] = np.where(west, f1w, f[1]); f[5] = np.where(west, f5w, f[5]); f[8] = np.where(west, f8w, f[8])
    f[3] = np.where(east, f3e, f[3]); f[7] = np.where(east, f7e, f[7]); f[6] = np.where(east, f6e, f[6])
    # moments (collision.cu:264-275)
    rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8]
    ux = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]) / rho
    uy = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]) / rho
    ex = ux + f32(Fx) / (rho * f32(ti)) if Fx != 0 else ux
    ey = uy + f32(Fy) / (rho * f32(ti)) if Fy != 0 else uy
    # BGK (collision.cu:297-330)
    xsq, ysq = ex * ex, ey * ey
    sq = xsq + ysq
    w0, w1, w5 = f32(16. / 36.) * rho, f32(4. / 36.) * rho, f32(1. / 36.) * rho
    u5, u6 = +ex + ey, -ex + ey
    t = f32(ti)
    eq = [w0 * (1. - 1.5 * sq),
          w1 * (1. + 3. * ex + 4.5 * xsq - 1.5 * sq),
          w1 * (1. + 3. * ey + 4.5 * ysq - 1.5 * sq),
          w1 * (1. - 3. * ex + 4.5 * xsq - 1.5 * sq),
          w1 * (1. - 3. * ey + 4.5 * ysq - 1.5 * sq),
          w5 * (1. + 3. * u5 + 4.5 * u5 * u5 - 1.5 * sq),
          w5 * (1. + 3. * u6 + 4.5 * u6 * u6 - 1.5 * sq),
          w5 * (1. - 3. * u5 + 4.5 * u5 * u5 - 1.5 * sq),
          w5 * (1. - 3. * u6 + 4.5 * u6 * u6 - 1.5 * sq)]
    out = np.stack([f[i] - t * (f[i] - eq[i]) for i in range(9)])
    fl = solid == 0
    return out, np.where(fl, rho, 0), np.where(fl, ux, 0), np.where(fl, uy, 0), fl


def ref_collide_d3q19(f, solid, ti, ux_in, rho_out, W):
    """collision.cu:357-541 on every node, vectorised."""
    f = [f[i].copy() for i in range(19)]
    for a, b in ((1, 2), (3, 4), (5, 6), (7, 12), (8, 11), (9, 14), (10, 13), (15, 18), (16, 17)):
        f[a], f[b] = f[b].copy(), f[a].copy()
    x = np.arange(f[0].size) % W
    ui, ro = f32(ux_in), f32(rho_out)
    west, east = x == 0, x == W - 1
    g = [a.copy() for a in f]
    # Zou-He velocity inlet (collision.cu:412-430); f9 / f10 use f11 where the
    # z closure needs f17 (the reference's slip, kept)
    rho_w = ((d(f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18])
              + 2. * d(f[2] + f[11] + f[12] + f[13] + f[14])) / (1. - f64(ui))).astype(f32)
    ru = d(rho_w * ui)
    ny = d(f[3] + f[15] + f[16] - f[4] - f[17] - f[18])
    nz = d(f[5] + f[11] + f[15] - f[6] - f[16] - f[18])
    w1 = (d(f[2]) + (1. / 3.) * ru).astype(f32)
    w8 = (d(f[11]) + (1. / 6.) * ru + 0.5 * ny).astype(f32)
    w7 = (d(f[12]) + (1. / 6.) * ru - 0.5 * ny).astype(f32)
    w9 = (d(f[14]) + (1. / 6.) * ru - 0.5 * nz).astype(f32)
    w10 = (d(f[13]) + (1. / 6.) * ru + 0.5 * nz).astype(f32)
    # Zou-He pressure outlet (collision.cu:433-451): f11 is overwritten first
    # and then read by the z closure of f14 / f13
    ux_e = (-1 + (d(f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18])
                  + 2. * d(f[1] + f[7] + f[8] + f[9] + f[10])) / f64(ro)).astype(f32)
    ru = d(ro * ux_e)
    e2 = (d(f[1]) - (1. / 3.) * ru).astype(f32)
    e11 = (d(f[8]) - (1. / 6.) * ru - 0.5 * ny).astype(f32)
    e12 = (d(f[7]) - (1. / 6.) * ru + 0.5 * ny).astype(f32)
    nz_e = d(f[5] + e11 + f[15] - f[6] - f[16] - f[18])
    e14 = (d(f[9]) - (1. / 6.) * ru + 0.5 * nz_e).astype(f32)
    e13 = (d(f[10]) - (1. / 6.) * ru - 0.5 * nz_e).astype(f32)
    for i, v in ((1, w1), (8, w8), (7, w7), (9, w9), (10, w10)):
        g[i] = np.where(west, v, g[i])
    for i, v in ((2, e2), (11, e11), (12, e12), (14, e14), (13, e13)):
        g[i] = np.where(east, v, g[i])
    f = g
    rho = f[0]
    for i in range(1, 19):
        rho = rho + f[i]
    ux = (f[1] + f[7] + f[8] + f[9] + f[10] - f[2] - f[11] - f[12] - f[13] - f[14]) / rho
    uy = (f[3] + f[7] + f[11] + f[15] + f[16] - f[4] - f[8] - f[12] - f[17] - f[18]) / rho
    uz = (f[5] + f[9] + f[13] + f[15] + f[17] - f[6] - f[10] - f[14] - f[16] - f[18]) / rho
    xsq, ysq, zsq = ux * ux, uy * uy, uz * uz
    sq = xsq + ysq + zsq
    w0, w1_, w7_ = f32(12. / 36.) * rho, f32(2. / 36.) * rho, f32(1. / 36.) * rho
    a7, a8, a9, a10, a15, a16 = ux + uy, ux - uy, ux + uz, ux - uz, uy + uz, uy - uz
    eq = [w0 * (1. - 1.5 * sq),
          w1_ * (1. + 3. * ux + 4.5 * xsq - 1.5 * sq), w1_ * (1. - 3. * ux + 4.5 * xsq - 1.5 * sq),
          w1_ * (1. + 3. * uy + 4.5 * ysq - 1.5 * sq), w1_ * (1. - 3. * uy + 4.5 * ysq - 1.5 * sq),
          w1_ * (1. + 3. * uz + 4.5 * zsq - 1.5 * sq), w1_ * (1. - 3. * uz + 4.5 * zsq - 1.5 * sq),
          w7_ * (1. + 3. * a7 + 4.5 * a7 * a7 - 1.5 * sq), w7_ * (1. + 3. * a8 + 4.5 * a8 * a8 - 1.5 * sq),
          w7_ * (1. + 3. * a9 + 4.5 * a9 * a9 - 1.5 * sq), w7_ * (1. + 3. * a10 + 4.5 * a10 * a10 - 1.5 * sq),
          w7_ * (1. - 3. * a8 + 4.5 * a8 * a8 - 1.5 * sq), w7_ * (1. - 3. * a7 + 4.5 * a7 * a7 - 1.5 * sq),
          w7_ * (1. - 3. * a10 + 4.5 * a10 * a10 - 1.5 * sq), w7_ * (1. - 3. * a9 + 4.5 * a9 * a9 - 1.5 * sq),
          w7_ * (1. + 3. * a15 + 4.5 * a15 * a15 - 1.5 * sq), w7_ * (1. + 3. * a16 + 4.5 * a16 * a16 - 1.5 * sq),
          w7_ * (1. - 3. * a16 + 4.5 * a16 * a16 - 1.5 * sq), w7_ * (1. - 3. * a15 + 4.5 * a15 * a15 - 1.5 * sq)]
    t = f32(ti)
    out = np.stack([f[i] - t * (f[i] - eq[i]) for i in range(19)])
    fl = solid == 0
    return out, np.where(fl, rho, 0), np.where(fl, ux, 0), np.where(fl, uy, 0), np.where(fl, uz, 0), fl


def _pops(rng, w, N, scale=0.08):
    return (w[:, None] * (1.0 + scale * rng.standard_normal((w.size, N)))).astype(f32)


def test_d2q9_collision_bits_match_reference_expressions():
    rng = np.random.default_rng(21)
    W, H = 16, 6
    N = W * H
    for ti, ux_in, rho_out, Fx, Fy in ((1.0 / 0.6, 0.05, 1.0, 0.0, 0.0), (1.25, 0.08, 1.02, 1e-5, -3e-5),
                                       (0.7, 0.02, 0.98, 0.0, 2e-4)):
        f = _pops(rng, O.W9F.astype(f64), N)
        solid = (rng.uniform(size=N) < 0.1).astype(np.uint8)
        want, wrho, wux, wuy, fl = ref_collide_d2q9(f, solid, ti, ux_in, rho_out, Fx, Fy, W)
        got = f.copy()
        ux, uy, rho, psi, ex, ey = (np.zeros(N, f32) for _ in range(6))
        plot = np.zeros(8 * N, f32)
        O.lib.orc_ref_collision_d2q9(O.P(plot), O.P(got), O.P(solid), O.P(ux), O.P(uy), O.P(rho), O.P(psi),
                                     O.P(ex), O.P(ey), W, H, 0, 1, O.W9F[0], O.W9F[1], O.W9F[5], f32(ti),
                                     ux_in, 0.0, rho_out, Fx, Fy, 0.0, 0.0, 1)
        np.testing.assert_array_equal(got[:, fl].view(np.uint32), want[:, fl].view(np.uint32))
        np.testing.assert_array_equal(got[:, ~fl], f[:, ~fl])          # solids untouched
        for a, b in ((rho, wrho), (ux, wux), (uy, wuy)):
            np.testing.assert_array_equal(a.view(np.uint32), b.astype(f32).view(np.uint32))
        pl = plot[:4 * N].reshape(N, 4)
        np.testing.assert_array_equal(pl[fl], np.stack([wux, wuy, 0 * wux, wrho], 1)[fl])


def test_d3q19_collision_bits_match_reference_expressions():
    rng = np.random.default_rng(22)
    W, H, D = 8, 4, 3
    N = W * H * D
    for ti, ux_in, rho_out in ((1.0 / 0.56, 0.05, 1.0), (1.1, 0.09, 1.03)):
        f = _pops(rng, O.W19F.astype(f64), N)
        solid = (rng.uniform(size=N) < 0.1).astype(np.uint8)
        want, wrho, wux, wuy, wuz, fl = ref_collide_d3q19(f, solid, ti, ux_in, rho_out, W)
        got = f.copy()
        ux, uy, uz, rho = (np.zeros(N, f32) for _ in range(4))
        plot = np.zeros(4 * N, f32)
        O.lib.orc_ref_collision_d3q19(O.P(plot), O.P(got), O.P(solid), O.P(ux), O.P(uy), O.P(uz), O.P(rho),
                                      W, H, D, 0, 1, O.W19F[0], O.W19F[1], O.W19F[7], f32(ti), ux_in, 0.0, 0.0,
                                      rho_out, 1)
        np.testing.assert_array_equal(got[:, fl].view(np.uint32), want[:, fl].view(np.uint32))
        for a, b in ((rho, wrho), (ux, wux), (uy, wuy), (uz, wuz)):
            np.testing.assert_array_equal(a.view(np.uint32), b.astype(f32).view(np.uint32))
