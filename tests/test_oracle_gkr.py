"""Pins of the oracle's Godunov-Kolgan-Rodionov step (Kernel_GKR,
kernel.cu:2663-3284; predictor kernel.cu:1611-1720; POTOK_Korolkov
Gas_Dinamic.cu:70-181).

* POTOK_Korolkov is the physical ideal-MHD flux: against the textbook flux
  (tests/textbook.py) for random states and every axis normal.
* The predictor is a step of the conserved variables: minus the divergence of
  the physical fluxes of the four half-slope face states, plus the
  axisymmetric source terms written from the cylindrical-form equations
  (no Powell term), then back to primitives.
* The whole step on interior cells of a random smooth state against a
  restatement built from those pieces and the textbook Riemann solvers
  (Miyoshi-Kusano HLLD, Li's HLLC-C): the minmod slopes (the vy slot taking
  the pressure slope), the five predictors at their radii, the
  predictor-averaged half-slope face states, the corrector from the mean of
  the cell and its predictor over dt/2.
* The frozen sphere: inside copied, outside bit for bit the step without it.
* The quiescent outer state (rho = 1, p = 1/gamma, v = B = 0) is steady.
What these leave open: the side rules of the first and last two rows and
columns (kernel.cu:2697-2957), which follow the source's text (including its
second-neighbour choices at n = 1 and m = 0).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import textbook as TB

SPI4 = 3.544907701811032
CPI4 = 4.0 * math.pi
CPI8 = 8.0 * math.pi
G53 = 5.0 / 3.0
PROD_BC = (O.BC_SYMMETRY, O.BC_OUTFLOW, O.BC_AXIS, O.BC_INFLOW)

FRAMES = {(1, 0, 0): ((0, 1, 0), (0, 0, 1)), (0, 1, 0): ((0, 0, 1), (1, 0, 0)), (0, 0, 1): ((1, 0, 0), (0, 1, 0)),
          (-1, 0, 0): ((0, 0, -1), (0, -1, 0)), (0, -1, 0): ((-1, 0, 0), (0, 0, -1)),
          (0, 0, -1): ((-1, 0, 0), (0, -1, 0))}


def phys_flux_lab(gamma, q, n):
    """textbook physical flux of q = (rho, p, vx, vy, vz, Bx, By, Bz), Gaussian,
    along the axis normal n, as (F_rho, F_m(3), F_B(3), F_E)"""
    n = tuple(int(v) for v in n)
    t, m = FRAMES[n]
    dot = lambda a, b: a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
    v, B = q[2:5], [x / SPI4 for x in q[5:8]]
    F = TB.mhd_phys_flux_hl(gamma, q[0], q[1], dot(v, n), dot(v, t), dot(v, m), dot(B, n), dot(B, t), dot(B, m))
    mom = [F[1] * n[i] + F[2] * t[i] + F[3] * m[i] for i in range(3)]
    bfl = [SPI4 * (F[4] * t[i] + F[5] * m[i]) for i in range(3)]
    return np.array([F[0], *mom, *bfl, F[6]])


def test_potok_is_the_physical_flux():
    rng = np.random.default_rng(1)
    for i in range(3000):
        q = (rng.uniform(0.2, 2), rng.uniform(0.1, 2), *rng.uniform(-1, 1, 3), *rng.uniform(-6, 6, 3))
        n = list(FRAMES)[i % 6]
        got = O.potok(G53, q, n)
        exp = phys_flux_lab(G53, q, n)
        assert np.max(np.abs(got - exp)) <= 1e-13 * (np.max(np.abs(exp)) + 1), (i, got, exp)


def cons(q, gamma):
    r, p, u, v, bx, by, bz = q
    return np.array([r, r * u, r * v, bx, by, bz,
                     p / (gamma - 1) + 0.5 * r * (u * u + v * v) + (bx * bx + by * by + bz * bz) / CPI8])


def prim(U, gamma):
    r, mu, mv, bx, by, bz, E = U
    u, v = mu / r, mv / r
    p = (gamma - 1) * (E - 0.5 * r * (u * u + v * v) - (bx * bx + by * by + bz * bz) / CPI8)
    return np.array([r, p, u, v, bx, by, bz])


def cyl_source_no_powell(q, y, gamma):
    """d/dt of (rho, rho u, rho v, Bx, By, Bz, E) from the 1/r terms of the
    axisymmetric (z = x, r = y) ideal-MHD equations, w = 0, Gaussian B:
    r-fluxes over r and the hoop stress Pi_rr - Pi_phiphi."""
    r, p, u, v, bx, by, bz = q
    b2 = bx * bx + by * by + bz * bz
    P = p + b2 / CPI8
    E = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v) + b2 / CPI8
    ub = u * bx + v * by
    return np.array([
        -r * v / y,
        -(r * u * v - bx * by / CPI4) / y,
        -(r * v * v + (bz * bz - by * by) / CPI4) / y,
        -(v * bx - u * by) / y,
        0.0,
        0.0,
        -((E + P) * v - by * ub / CPI4) / y,
    ])


def predictor_ref(gamma, dx, dy, Q, DX, DY, st, y):
    """conserved step of st: divergence of the physical fluxes of the four
    half-slope face states (rho, p falling back to the cell value) plus the
    cylindrical sources, back to primitives (no clamp fires here)"""
    div = np.zeros(7)
    for D, sgn, n, ln in ((DX, -1, (-1, 0, 0), dy), (DX, 1, (1, 0, 0), dy), (DY, 1, (0, 1, 0), dx),
                          (DY, -1, (0, -1, 0), dx)):
        f = np.array(Q) + sgn * np.array(D) / 2.0
        if f[0] <= 0:
            f[0] = Q[0]
        if f[1] <= 0:
            f[1] = Q[1]
        F = phys_flux_lab(gamma, (f[0], f[1], f[2], f[3], 0.0, f[4], f[5], f[6]), n)
        # (rho, mx, my, Bx, By, Bz, E) order
        div += ln * np.array([F[0], F[1], F[2], F[4], F[5], F[6], F[7]])
    U = cons(Q, gamma) - st * div / (dx * dy) + st * cyl_source_no_powell(Q, y, gamma)
    return prim(U, gamma)


def _grid(nx, ny, r_frozen=0.0, y0=None, dx=0.05, dy=0.04):
    return O.make_grid(nx, ny, dx=dx, dy=dy, x0=dx / 2, y0=dy / 2 if y0 is None else y0, gamma=G53, krit=0.25,
                       bc=PROD_BC, r_frozen=r_frozen, axisym=True)


def test_predictor_is_a_conserved_step_with_axisymmetric_sources():
    rng = np.random.default_rng(2)
    g = _grid(4, 4)
    for i in range(500):
        Q = np.array([rng.uniform(0.5, 2), rng.uniform(0.5, 2), *rng.uniform(-1, 1, 2), *rng.uniform(-3, 3, 3)])
        DX = rng.uniform(-0.05, 0.05, 7)
        DY = rng.uniform(-0.05, 0.05, 7)
        if i % 10 == 0:
            DX[0] = 3.0 * Q[0]  # rho falls back to the cell value on the west face
        y = rng.uniform(0.3, 3.0)
        st = rng.uniform(1e-4, 1e-3)
        got = O.gkr_predictor(g, Q, DX, DY, st, y)
        exp = predictor_ref(G53, g.dx, g.dy, Q, DX, DY, st, y)
        assert np.max(np.abs(got - exp) / (np.abs(exp) + 1e-3)) <= 1e-11, (i, got, exp)


def minmod(a, b):
    return np.where(a * b > 0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)), 0.0)


def textbook_face_lab(metod, gamma, L, R, n2):
    """textbook HLLD (metod 3) / HLLC-C (metod 2) flux through an in-plane axis
    face, lab frame, (F_rho, F_m(3), F_B(3), F_E), with the normal-field
    dissipation -max(|SL|, |SR|)(Bn_R - Bn_L) (pinned in test_oracle_flux)"""
    n = (int(n2[0]), int(n2[1]), 0)
    t, m = FRAMES[n]
    dot = lambda a, b: a[0] * b[0] + a[1] * b[1] + a[2] * b[2]

    def face(q):
        v, B = (q[2], q[3], 0.0), [x / SPI4 for x in q[4:7]]
        return (q[0], q[1], dot(v, n), dot(v, t), dot(v, m), dot(B, n), dot(B, t), dot(B, m))

    ql, qr = face(L), face(R)
    if metod == 2:
        F8, (SL, SR) = TB.li_hllc(gamma, ql, qr)
        F8 = list(F8)
    else:
        F7, (SL, SR) = TB.mk_hlld(gamma, ql, qr, deg_abs=1e-5, alfven_abs=1e-6)
        F8 = [F7[0], F7[1], F7[2], F7[3], 0.0, F7[4], F7[5], F7[6]]
    F8[4] = -max(abs(SL), abs(SR)) * (qr[5] - ql[5])
    mom = [F8[1] * n[i] + F8[2] * t[i] + F8[3] * m[i] for i in range(3)]
    bfl = [SPI4 * (F8[4] * n[i] + F8[5] * t[i] + F8[6] * m[i]) for i in range(3)]
    return np.array([F8[0], *mom, *bfl, F8[7]]), max(abs(SL), abs(SR))


def gkr_interior_ref(g, q, metod, Tdo, n, m):
    """the step of interior cell (n, m) (at least two cells from every side)
    from the textbook pieces; q[m, n] = (rho, p, vx, vy, Bx, By, Bz)"""
    dx, dy, gam = g.dx, g.dy, g.gamma
    src = [0, 1, 2, 1, 4, 5, 6]  # the vy slot takes the pressure slope

    def sx(a, b):  # slope along x at cell (a, b) from its two x-neighbours
        return np.array([dx * minmod((q[b, a + 1, c] - q[b, a, c]) / dx, (q[b, a, c] - q[b, a - 1, c]) / dx)
                         for c in src])

    def sy(a, b):
        return np.array([dy * minmod((q[b + 1, a, c] - q[b, a, c]) / dy, (q[b, a, c] - q[b - 1, a, c]) / dy)
                         for c in src])

    y = g.y0 + m * dy
    cells = {1: (n, m, y), 2: (n + 1, m, y), 3: (n, m - 1, y - dy), 4: (n - 1, m, y), 5: (n, m + 1, y + dy)}
    Qb, D = {}, {}
    for k, (a, b, yk) in cells.items():
        DX, DY = sx(a, b), sy(a, b)
        QQ = predictor_ref(gam, dx, dy, q[b, a], DX, DY, Tdo / 2.0, yk)
        Qb[k] = (q[b, a] + QQ) / 2.0
        D[k] = (DX, DY)
    div = np.zeros(7)
    tmin = 1e300
    # east, south, west, north
    for k, din, dnb, sgn, nrm, rad, ln in ((2, D[1][0], D[2][0], 1, (1, 0), dx, dy),
                                           (3, D[1][1], D[3][1], -1, (0, -1), dy, dx),
                                           (4, D[1][0], D[4][0], -1, (-1, 0), dx, dy),
                                           (5, D[1][1], D[5][1], 1, (0, 1), dy, dx)):
        L = Qb[1] + sgn * din / 2.0
        R = Qb[k] - sgn * dnb / 2.0
        for c in (0, 1):
            if L[c] <= 0:
                L[c] = Qb[1][c]
            if R[c] <= 0:
                R[c] = Qb[k][c]
        F, S = textbook_face_lab(metod, gam, L, R, nrm)
        tmin = min(tmin, g.krit * rad / S)
        div += ln * np.array([F[0], F[1], F[2], F[4], F[5], F[6], F[7]])
    h = Tdo / 2.0
    U = cons(Qb[1], gam) - h * div / (dx * dy) + h * cyl_source_no_powell(Qb[1], y, gam)
    return prim(U, gam), tmin


def random_smooth(g, seed):
    rng = np.random.default_rng(seed)
    x = g.x0 + np.arange(g.nx) * g.dx
    y = g.y0 + np.arange(g.ny) * g.dy
    X, Y = np.meshgrid(x, y)

    def field(lo, hi):
        a, b, c, d = rng.uniform(0.5, 3, 4)
        return lo + (hi - lo) * (0.5 + 0.25 * np.sin(a * X + b) + 0.25 * np.cos(c * Y + d))

    s = np.stack([field(0.5, 1.5), field(0.5, 1.5)], -1)
    u = np.stack([field(-0.5, 0.5), field(-0.5, 0.5)], -1)
    b = np.stack([field(-2, 2), field(-2, 2), field(-2, 2)], -1)
    return np.ascontiguousarray(s), np.ascontiguousarray(u), np.ascontiguousarray(b)


@pytest.mark.parametrize("metod", [3, 2])
def test_interior_step_matches_the_textbook_scheme(metod):
    g = _grid(12, 11, y0=0.7)
    s, u, b = random_smooth(g, 3 + metod)
    Tdo = 2e-3
    s2, u2, b2, T = O.ref_gkr_step(g, s, u, b, 1e7, Tdo, metod)
    q = np.concatenate([s, u, b], -1)
    tmin = 1e300
    for m in range(2, g.ny - 2):
        for n in range(2, g.nx - 2):
            exp, t = gkr_interior_ref(g, q, metod, Tdo, n, m)
            got = np.array([*s2[m, n], *u2[m, n], *b2[m, n]])
            assert np.max(np.abs(got - exp) / (np.abs(exp) + 1e-2)) <= 1e-11, (n, m, got, exp)
            tmin = min(tmin, t)
    assert T <= tmin * (1 + 1e-14)  # the global minimum includes the boundary cells


def test_frozen_sphere_is_copied_and_the_rest_unchanged():
    g0 = _grid(30, 20)
    g1 = _grid(30, 20, r_frozen=0.45)
    s, u, b = random_smooth(g0, 9)
    r0 = O.ref_gkr_step(g0, s, u, b, 1e7, 1e-3, 3)
    r1 = O.ref_gkr_step(g1, s, u, b, 1e7, 1e-3, 3)
    x = g1.x0 + np.arange(g1.nx) * g1.dx
    y = g1.y0 + np.arange(g1.ny) * g1.dy
    X, Y = np.meshgrid(x, y)
    inside = np.sqrt(X * X + Y * Y) < g1.r_frozen
    assert 20 < inside.sum() < inside.size
    for a0, a1, a in zip(r0[:3], r1[:3], (s, u, b)):
        np.testing.assert_array_equal(a1[inside], a[inside])
        np.testing.assert_array_equal(a1[~inside], a0[~inside])


def test_quiescent_state_is_steady():
    g = _grid(15, 12)
    s = np.zeros((g.ny, g.nx, 2))
    s[..., 0] = 1.0
    s[..., 1] = 1.0 / G53
    u = np.zeros((g.ny, g.nx, 2))
    b = np.zeros((g.ny, g.nx, 3))
    s2, u2, b2, _ = O.ref_gkr_step(g, s, u, b, 1e7, 1e-2, 3)
    np.testing.assert_allclose(s2, s, rtol=0, atol=1e-15)
    np.testing.assert_allclose(u2, u, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(b2, b)
