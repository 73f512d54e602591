"""Pins of the oracle's D2Q9 Shan-Chen / adhesion pass (orc_ref_collision_d2q9
with G < 0, the restatement of collision.cu:35-166 and the psi store at
collision.cu:291-295) by checks that do not restate its code:

* the per-node force against the textbook pseudo-potential force
  F(x) = -G psi(x) sum_n w_n psi(x + e_n) e_n evaluated with array shifts in
  the physical y-up frame (a transposed neighbour, a wrong sign or weight of
  any of the twelve stencil terms, or a wrong psi constant fails it);
* total momentum conservation on a periodic box (the pair forces cancel);
* the closed form of cohesion + adhesion next to solids for uniform psi:
  F = (G psi - Gads) psi sum_{n: x+e_n solid} w_n e_n  (solid psi = 0);
* flat-interface coexistence densities of psi = 4 exp(-200/rho), G = -120
  against the Shan-Chen mechanical-stability integral, and against the
  coexistence pair the reference quotes for this potential (85.7 / 524.39,
  simulation_d2q9.cu:150-153).

The force is read from the collision itself: with populations at rest
equilibrium f_i = w_i rho and tau = 1 the post-collision first moment is
rho * ueq = F exactly in exact arithmetic (collision.cu:95-96, 141-164)."""
import math

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.optimize import brentq

import oracle as O

# storage rows grow downward; physical c_y = -storage c_y
C9S = np.array([[0, 0], [1, 0], [0, -1], [-1, 0], [0, 1], [1, -1], [-1, -1], [-1, 1], [1, 1]])
C9P = np.stack([C9S[:, 0], -C9S[:, 1]], 1)
W9 = np.array([16 / 36] + [4 / 36] * 4 + [1 / 36] * 4)
G_ST, PSI0, RHO0 = -120.0, 4.0, 200.0


def psi_st(rho):
    return PSI0 * np.exp(-RHO0 / np.asarray(rho, np.float64))


def neighbour(a, e):
    """a at x + e (physical e, periodic); arrays are [row, x] with rows down."""
    return np.roll(np.roll(a, -int(e[0]), axis=1), int(e[1]), axis=0)


def collide_rest(rho, solid, G, Gads, tau_inv=1.0):
    """One call of the oracle's collision (with the Shan-Chen pass) on rest
    populations f_i = w_i rho; returns the physical first moment after it."""
    H, W = rho.shape
    N = W * H
    f = (O.W9F.astype(np.float32)[:, None, None] * rho.astype(np.float32)[None]).astype(np.float32)
    f[:, solid.astype(bool)] = 0.0
    sol = np.ascontiguousarray(solid.astype(np.uint8))
    ux, uy, rh, psi, ex, ey = (np.zeros(N, np.float32) for _ in range(6))
    plot = np.zeros(8 * N, np.float32)
    O.lib.orc_ref_collision_d2q9(O.P(plot), O.P(f), O.P(sol), O.P(ux), O.P(uy), O.P(rh), O.P(psi), O.P(ex),
                                 O.P(ey), W, H, 0, 1, O.W9F[0], O.W9F[1], O.W9F[5], tau_inv, 0.0, 0.0, 1.0,
                                 0.0, 0.0, G, Gads, 0)
    fd = f.astype(np.float64)
    px = np.tensordot(C9P[:, 0].astype(np.float64), fd, 1)
    py = np.tensordot(C9P[:, 1].astype(np.float64), fd, 1)
    return px, py, psi.reshape(H, W)


def textbook_force(rho, solid, G, Gads):
    fluid = ~solid.astype(bool)
    psi = np.where(fluid, psi_st(rho), 0.0)
    sx = np.zeros_like(psi)
    sy = np.zeros_like(psi)
    ax = np.zeros_like(psi)
    ay = np.zeros_like(psi)
    for n in range(1, 9):
        e = C9P[n]
        sx += W9[n] * neighbour(psi, e) * e[0]
        sy += W9[n] * neighbour(psi, e) * e[1]
        s = neighbour(solid.astype(np.float64), e)
        ax += W9[n] * s * e[0]
        ay += W9[n] * s * e[1]
    fx = -G * psi * sx - Gads * psi * ax
    fy = -G * psi * sy - Gads * psi * ay
    return np.where(fluid, fx, 0.0), np.where(fluid, fy, 0.0), psi


def test_psi_store_is_the_sukop_thorne_potential():
    rho = np.array([[50.0, 85.7, 200.0, 524.39], [1000.0, 120.0, 300.0, 80.0]])
    _, _, psi = collide_rest(rho, np.zeros(rho.shape, np.uint8), G_ST, 0.0)
    # rho is re-summed from nine FP32 populations first: a few ulp, amplified
    # by dln(psi)/dln(rho) = 200/rho <= 4 here
    np.testing.assert_allclose(psi, psi_st(rho.astype(np.float32)), rtol=3e-6)


def test_force_matches_textbook_stencil_on_a_random_field():
    rng = np.random.default_rng(3)
    H, W = 12, 16
    rho = rng.uniform(80.0, 520.0, (H, W))
    solid = np.zeros((H, W), np.uint8)
    px, py, _ = collide_rest(rho, solid, G_ST, 0.0)
    fx, fy, psi = textbook_force(rho.astype(np.float32), solid, G_ST, 0.0)
    scale = np.abs(G_ST) * psi.max() ** 2 * W9[5]     # one diagonal stencil term
    assert np.abs(fx).max() > 2 * scale and np.abs(fy).max() > 2 * scale
    # FP32 populations of size ~ w rho: rounding ~1e-4 absolute per node
    np.testing.assert_allclose(px.reshape(H, W), fx, atol=1e-3 * scale)
    np.testing.assert_allclose(py.reshape(H, W), fy, atol=1e-3 * scale)


def test_cohesion_and_adhesion_closed_form_next_to_solids():
    """Uniform rho beside a solid floor row and a solid block: solid psi is 0,
    so F = (G psi - Gads) psi sum_{solid n} w_n e_n at every fluid node."""
    H, W = 16, 16
    solid = np.zeros((H, W), np.uint8)
    solid[H - 1, :] = 1                      # floor (bottom storage row)
    solid[5:8, 6:10] = 1                     # a block: faces and corners
    fluid = solid == 0
    for rho0 in (120.0, 200.0, 480.0):
        rho = np.full((H, W), rho0)
        psi = float(psi_st(np.float32(rho0)))
        sw = np.zeros((2, H, W))
        for n in range(1, 9):
            s = neighbour(solid.astype(np.float64), C9P[n])
            sw[0] += W9[n] * s * C9P[n][0]
            sw[1] += W9[n] * s * C9P[n][1]
        assert (np.abs(sw[1][fluid]) > 0).sum() >= 2 * W
        for Gads in (0.0, -187.16, -327.79, -46.534):
            px, py, _ = collide_rest(rho, solid, G_ST, Gads)
            want = (G_ST * psi - Gads) * psi
            tol = 2e-5 * abs(G_ST) * psi * psi + 1e-4
            np.testing.assert_allclose(px.reshape(H, W)[fluid], want * sw[0][fluid], rtol=1e-5, atol=tol)
            np.testing.assert_allclose(py.reshape(H, W)[fluid], want * sw[1][fluid], rtol=1e-5, atol=tol)
        # nodes away from solids feel nothing
        far = fluid & (np.abs(sw).sum(0) == 0)
        assert far.sum() > 50


def _ref_run(f, solid, steps, G=G_ST, Gads=0.0, tau_inv=1.0):
    H, W = solid.shape
    N = W * H
    sol = np.ascontiguousarray(solid.astype(np.uint8).ravel())
    sb = np.zeros(N, np.uint8)
    O.lib.orc_ref_solid_bound_d2q9(O.P(sol), O.P(sb), W, H)
    ux, uy, rh, psi, ex, ey = (np.zeros(N, np.float32) for _ in range(6))
    plot = np.zeros(8 * N, np.float32)
    vc = np.zeros(4, np.float32)
    O.lib.orc_ref_run_d2q9(1, W, H, O.P(f), O.P(sol), O.P(sb), O.P(ux), O.P(uy), O.P(rh), O.P(psi), O.P(ex),
                           O.P(ey), O.P(plot), 0.0, 0.0, 1.0, tau_inv, 0.0, 0.0, G, Gads, int(steps), 0, 0,
                           O.P(vc), 0)
    return rh.reshape(H, W), ux.reshape(H, W), uy.reshape(H, W)


def test_total_momentum_conserved_on_a_periodic_box():
    rng = np.random.default_rng(11)
    H, W = 32, 24
    rho = rng.uniform(150.0, 400.0, (H, W))
    f = np.ascontiguousarray((O.W9F[:, None, None] * rho.astype(np.float32)[None]).astype(np.float32))
    mass0 = f.astype(np.float64).sum()
    # the sum of |F| over nodes: what an unbalanced stencil would leave behind
    fx, fy, _ = textbook_force(rho.astype(np.float32), np.zeros((H, W), np.uint8), G_ST, 0.0)
    fabs = np.abs(fx).sum() + np.abs(fy).sum()
    for _ in range(5):
        _ref_run(f, np.zeros((H, W), np.uint8), 10)
        fd = f.astype(np.float64)
        p = [np.tensordot(C9P[:, a].astype(np.float64), fd, 1).sum() for a in (0, 1)]
        assert abs(p[0]) < 1e-5 * fabs and abs(p[1]) < 1e-5 * fabs, (p, fabs)
        assert abs(fd.sum() - mass0) < 1e-6 * mass0


def _coexistence_sukop_thorne(G=G_ST):
    """Gas / liquid densities from the mechanical-stability condition
    int_{rg}^{rl} (p0 - p(rho)) psi'/psi^2 drho = 0 with the D2Q9 equation of
    state p = rho/3 + G psi^2 / 6 (Shan & Chen 1993, Shan 2008)."""
    psi = lambda r: PSI0 * math.exp(-RHO0 / r)
    dpsi = lambda r: psi(r) * RHO0 / r ** 2
    p = lambda r: r / 3 + G * psi(r) ** 2 / 6
    dp = lambda r: 1 / 3 + G * psi(r) * dpsi(r) / 3
    rs = np.linspace(20.0, 2000.0, 20000)
    d = np.array([dp(r) for r in rs])
    idx = np.where(np.diff(np.sign(d)))[0]
    r1 = brentq(dp, rs[idx[0]], rs[idx[0] + 1])
    r2 = brentq(dp, rs[idx[1]], rs[idx[1] + 1])

    def roots(p0):
        return brentq(lambda r: p(r) - p0, 1.0, r1), brentq(lambda r: p(r) - p0, r2, 5000.0)

    def area(p0):
        rg, rl = roots(p0)
        return quad(lambda r: (p0 - p(r)) * dpsi(r) / psi(r) ** 2, rg, rl, limit=400)[0]

    p0 = brentq(area, max(p(r2), p(1.0)) + 1e-9, p(r1) - 1e-9)
    return roots(p0)


def test_mechanical_stability_near_the_reference_pair():
    rg, rl = _coexistence_sukop_thorne()
    # simulation_d2q9.cu:150-153 seeds bubbles with 85.7 / 524.39; the
    # continuum integral sits a few percent off the lattice values
    assert abs(rl - 524.39) / 524.39 < 0.01 and abs(rg - 85.7) / 85.7 < 0.04, (rg, rl)


def test_flat_interface_coexistence():
    H, W = 96, 4
    y = np.arange(H)
    rho = np.where((y >= H // 4) & (y < 3 * H // 4), 480.0, 110.0)[:, None] * np.ones((1, W))
    f = np.ascontiguousarray((O.W9F[:, None, None] * rho.astype(np.float32)[None]).astype(np.float32))
    r1, _, _ = _ref_run(f, np.zeros((H, W), np.uint8), 12000)
    r2, ux, _ = _ref_run(f, np.zeros((H, W), np.uint8), 4000)
    prof = r2[:, 0]
    assert np.ptp(r2, axis=1).max() < 1e-3 and np.abs(ux).max() == 0.0    # stays a flat interface
    assert np.abs(r2 - r1).max() < 1e-3 * prof.max()                        # steady
    # the reference's coexistence pair for this potential (simulation_d2q9.cu:150-153)
    assert abs(prof.max() - 524.39) / 524.39 < 2e-3, prof.max()
    assert abs(prof.min() - 85.7) / 85.7 < 2e-3, prof.min()
    # and the mechanical-stability integral, to its known lattice offset
    rg, rl = _coexistence_sukop_thorne()
    assert abs(prof.max() - rl) / rl < 0.01, (prof.max(), rl)
    assert abs(prof.min() - rg) / rg < 0.04, (prof.min(), rg)
