"""Pins of the oracle's production options of the reference-layout steps:
MHD add2 / Kernel_TVD (axisymmetric source terms, the four production side
kinds, the frozen inner sphere; kernel.cu:1918-2190, 2218-2650) and Euler
Kernel_TVD_old (axisymmetric source terms, inner sphere; kernel.cu:3303,
3651-3662).

* Sources: on a uniform state every face flux cancels, so one step is a
  forward-Euler step of the geometric and Powell source terms alone. The
  expected update is written from the axisymmetric ideal-MHD equations in
  cylindrical (z = x, r = y) form: stress tensor Pi_rr - Pi_phiphi hoop term,
  r-flux of each conserved field over r, Powell's 8-wave term with
  div B = dBz/dz + dBr/dr + Br/r.
* Side kinds: the step on a grid with the production sides equals, bit for
  bit, the interior of the same step on a grid two cells larger on every
  side whose extra cells hold the ghost states explicitly, filled here from
  the reference's rules (symmetry west, outflow east, axis south, inflow
  north).
* Frozen sphere: cells inside keep their state, every other cell is the
  step without the sphere bit for bit.
* The quiescent outer state of the production run is a steady state.
* Rows below the axis: add2 evaluates them in the mirrored frame
  (kernel.cu:1885-1889, 2026-2057, 2177-2182), so its step on the mirror
  image of a grid above the axis is the mirror image of that grid's step,
  bit for bit (coordinates exact in binary); Kernel_TVD and Kernel_TVD_old
  divide by the signed y, which keeps the same mirror relation to rounding.
"""
import numpy as np
import pytest

from golikov_2d_amd import problems as PB
from oracle import oracle as O

CPI4 = 4.0 * np.pi
CPI8 = 8.0 * np.pi
G53 = 5.0 / 3.0
PROD_BC = (O.BC_SYMMETRY, O.BC_OUTFLOW, O.BC_AXIS, O.BC_INFLOW)


def cyl_source(q, y, gamma):
    """d/dt of (rho, rho u, rho v, Bx, By, Bz, E) for a uniform state at
    radius y from the cylindrical-form equations (w = 0)."""
    r, p, u, v, bx, by, bz = q
    b2 = bx * bx + by * by + bz * bz
    P = p + b2 / CPI8
    E = p / (gamma - 1.0) + 0.5 * r * (u * u + v * v) + b2 / CPI8
    pi_rr = r * v * v + P - by * by / CPI4
    pi_pp = P - bz * bz / CPI4
    divB = by / y                      # only the Br / r part is non-zero
    ub = u * bx + v * by
    return np.array([
        -r * v / y,
        -(r * u * v - bx * by / CPI4) / y - divB * bx / CPI4,
        -(pi_rr - pi_pp) / y - divB * by / CPI4,
        -(v * bx - u * by) / y - divB * u,
        -divB * v,
        0.0,
        -((E + P) * v - by * ub / CPI4) / y - divB * ub / CPI4,
    ])


@pytest.mark.parametrize("kind", ["add2", "kernel_tvd"])
def test_axisymmetric_sources_on_uniform_state(kind):
    q = (1.3, 0.9, 0.4, -0.3, 0.5, 0.7, -0.6)
    nx, ny, dx, dy, y0 = 6, 5, 0.1, 0.1, 0.45
    g = O.make_grid(nx, ny, dx=dx, dy=dy, y0=y0, gamma=G53, axisym=True)
    s = np.empty((ny, nx, 2)); u = np.empty((ny, nx, 2)); b = np.empty((ny, nx, 3))
    s[...] = (q[0], q[1]); u[...] = (q[2], q[3]); b[...] = q[4:]
    dt = 1e-3
    s2, u2, b2, _ = O.ref_mhd_step(kind, g, s, u, b, 1e7, dt, 0, 3)
    r, p = q[0], q[1]
    U = np.array([r, r * q[2], r * q[3], q[4], q[5], q[6],
                  p / (G53 - 1) + 0.5 * r * (q[2] ** 2 + q[3] ** 2) + (q[4] ** 2 + q[5] ** 2 + q[6] ** 2) / CPI8])
    for j in range(ny):
        y = y0 + j * dy
        W = U + dt * cyl_source(q, y, G53)
        ux, uy = W[1] / W[0], W[2] / W[0]
        pe = (W[6] - 0.5 * W[0] * (ux * ux + uy * uy) - (W[3] ** 2 + W[4] ** 2 + W[5] ** 2) / CPI8) * (G53 - 1)
        want = np.array([W[0], pe, ux, uy, W[3], W[4], W[5]])
        for i in range(nx):
            got = np.array([s2[j, i, 0], s2[j, i, 1], u2[j, i, 0], u2[j, i, 1], *b2[j, i]])
            assert np.allclose(got, want, rtol=1e-12, atol=1e-14), (j, i, got - want)
    # without the option the uniform state does not move
    g.axisym = 0
    s2, u2, b2, _ = O.ref_mhd_step(kind, g, s, u, b, 1e7, dt, 0, 3)
    assert np.allclose(s2, s, rtol=1e-14) and np.allclose(u2, u, atol=1e-14) and np.allclose(b2, b, atol=1e-14)


def smooth_state(g, seed):
    """Smooth state with every primitive non-zero and vx < 0 somewhere on the
    east edge (so the outflow rule acts)."""
    x = g.x0 + np.arange(g.nx) * g.dx
    y = g.y0 + np.arange(g.ny) * g.dy
    X, Y = np.meshgrid(x, y)
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 1.5, 8)
    s = np.empty((g.ny, g.nx, 2)); u = np.empty((g.ny, g.nx, 2)); b = np.empty((g.ny, g.nx, 3))
    s[..., 0] = 1.0 + 0.3 * np.sin(a[0] * X + 0.3) * np.cos(a[1] * Y)
    s[..., 1] = 0.8 + 0.2 * np.cos(a[2] * X - a[3] * Y)
    u[..., 0] = 0.5 * np.sin(a[4] * Y + 0.7) - 0.1
    u[..., 1] = 0.4 * np.cos(a[5] * X)
    b[..., 0] = 0.6 * np.cos(a[6] * Y)
    b[..., 1] = 0.5 * np.sin(a[7] * X + 0.2)
    b[..., 2] = 0.4 * np.cos(X + Y)
    return s, u, b


def explicit_ghosts(s, u, b, gamma, inflow_b_zero):
    """Copy of (s, u, b) two cells larger on each side, the extra cells set by
    the reference's production ghost rules (kernel.cu:1938-2016,
    2227-2405); corner cells are not read by the cross stencil."""
    ny, nx = s.shape[:2]
    S = np.zeros((ny + 4, nx + 4, 2)); Uv = np.zeros((ny + 4, nx + 4, 2)); B = np.zeros((ny + 4, nx + 4, 3))
    S[...] = (1.0, 1.0 / gamma)
    S[2:-2, 2:-2], Uv[2:-2, 2:-2], B[2:-2, 2:-2] = s, u, b
    for layer in (1, 2):
        # west symmetry: mirror, vx, Bx, By negated
        src = layer - 1
        S[2:-2, 2 - layer] = s[:, src]
        Uv[2:-2, 2 - layer] = u[:, src] * (-1, 1)
        B[2:-2, 2 - layer] = b[:, src] * (-1, -1, 1)
        # east outflow: edge cell, vx < 0 raised to 0.01
        S[2:-2, nx + 1 + layer] = s[:, -1]
        ue = u[:, -1].copy()
        ue[ue[:, 0] < 0, 0] = 0.01
        Uv[2:-2, nx + 1 + layer] = ue
        B[2:-2, nx + 1 + layer] = b[:, -1]
        # south axis: mirror, vy and Bz negated
        S[2 - layer, 2:-2] = s[src]
        Uv[2 - layer, 2:-2] = u[src] * (1, -1)
        B[2 - layer, 2:-2] = b[src] * (1, 1, -1)
        # north inflow: rho = 1, p = 1/gamma, v = 0, B of the edge cell or 0
        S[ny + 1 + layer, 2:-2] = (1.0, 1.0 / gamma)
        Uv[ny + 1 + layer, 2:-2] = 0.0
        B[ny + 1 + layer, 2:-2] = 0.0 if inflow_b_zero else b[-1]
    return S, Uv, B


@pytest.mark.parametrize("kind", ["add2", "kernel_tvd"])
@pytest.mark.parametrize("method", [2, 3])
@pytest.mark.parametrize("axisym", [False, True])
def test_production_sides_equal_explicit_ghost_cells(kind, method, axisym):
    nx, ny, h = 13, 11, 0.125
    x0, y0 = 0.0625, 3.0625                # exact in binary on both grids
    g = O.make_grid(nx, ny, dx=h, dy=h, x0=x0, y0=y0, gamma=G53, bc=PROD_BC, axisym=axisym)
    s, u, b = smooth_state(g, seed=nx + method)
    dt = 2e-3
    s2, u2, b2, _ = O.ref_mhd_step(kind, g, s, u, b, 1e7, dt, 0, method)
    S, Uv, B = explicit_ghosts(s, u, b, G53, inflow_b_zero=(kind == "kernel_tvd"))
    gb = O.make_grid(nx + 4, ny + 4, dx=h, dy=h, x0=x0 - 2 * h, y0=y0 - 2 * h, gamma=G53, axisym=axisym)
    S2, U2, B2, _ = O.ref_mhd_step(kind, gb, S, Uv, B, 1e7, dt, 0, method)
    np.testing.assert_array_equal(s2, S2[2:-2, 2:-2])
    np.testing.assert_array_equal(u2, U2[2:-2, 2:-2])
    np.testing.assert_array_equal(b2, B2[2:-2, 2:-2])
    # the side kinds act: the step differs from the transmissive one at the edges
    gt = O.make_grid(nx, ny, dx=h, dy=h, x0=x0, y0=y0, gamma=G53, axisym=axisym)
    s2t, _, b2t, _ = O.ref_mhd_step(kind, gt, s, u, b, 1e7, dt, 0, method)
    for sl in (np.s_[:, 0], np.s_[:, -1], np.s_[0, :], np.s_[-1, :]):
        assert np.any(b2t[sl] != b2[sl]) or np.any(s2t[sl] != s2[sl])


@pytest.mark.parametrize("kind", ["add2", "kernel_tvd"])
def test_frozen_sphere_keeps_state_and_leaves_the_rest(kind):
    nx, ny = 40, 24
    kw = PB.wind_grid_args(nx, ny, x_max=2.0, y_max=1.2)
    g = O.make_grid(nx, ny, **kw)
    s, u, b = smooth_state(g, seed=3)
    x = g.x0 + np.arange(nx) * g.dx
    y = g.y0 + np.arange(ny) * g.dy
    X, Y = np.meshgrid(x, y)
    d = np.sqrt(X * X + Y * Y)
    assert np.min(np.abs(d - g.r_frozen)) > 1e-6
    inside = d < g.r_frozen
    assert 5 < inside.sum() < inside.size // 4
    s2, u2, b2, T = O.ref_mhd_step(kind, g, s, u, b, 1e7, 1e-3, 0, 3)
    g.r_frozen = 0.0
    s2n, u2n, b2n, Tn = O.ref_mhd_step(kind, g, s, u, b, 1e7, 1e-3, 0, 3)
    np.testing.assert_array_equal(s2[inside], s[inside])
    np.testing.assert_array_equal(b2[inside], b[inside])
    np.testing.assert_array_equal(u2[~inside], u2n[~inside])
    np.testing.assert_array_equal(b2[~inside], b2n[~inside])
    assert np.any(s2n[inside] != s[inside])
    assert T >= Tn


@pytest.mark.parametrize("kind", ["add2", "kernel_tvd"])
def test_quiescent_outer_state_is_steady(kind):
    g = O.make_grid(30, 20, **PB.wind_grid_args(30, 20, x_max=3.0, y_max=2.0))
    g.r_frozen = 0.0
    s = np.empty((20, 30, 2)); s[...] = (1.0, 1.0 / G53)
    u = np.zeros((20, 30, 2)); b = np.zeros((20, 30, 3))
    s2, u2, b2, _ = O.ref_mhd_step(kind, g, s, u, b, 1e7, 1e-3, 0, 2)
    np.testing.assert_array_equal(s2, s)
    np.testing.assert_array_equal(u2, u)
    np.testing.assert_array_equal(b2, b)


def test_wind_initial_state_matches_closed_form():
    g = O.make_grid(161, 90, **PB.wind_grid_args(161, 90))
    s, u, b = PB.stellar_wind_prim(g)
    x = g.x0 + np.arange(g.nx) * g.dx
    y = g.y0 + np.arange(g.ny) * g.dy
    X, Y = np.meshgrid(x, y)
    d = np.hypot(X, Y)
    inner = d < 0.303
    assert inner.sum() > 10
    # mass flux rho |v| r^2 is constant in the wind, p rho^-gamma too
    flux = s[inner, 0] * np.hypot(u[inner, 0], u[inner, 1]) * d[inner] ** 2
    assert np.allclose(flux, flux[0], rtol=1e-13)
    ent = s[inner, 1] / s[inner, 0] ** G53
    assert np.allclose(ent, ent[0], rtol=1e-12)
    # B_phi r / sin(theta) constant, Alfven Mach number M_alf at r = rr_0
    bphi = b[inner, 2] * d[inner] / (Y[inner] / d[inner])
    assert np.allclose(bphi, np.sqrt(4 * np.pi) / PB.WIND_MALF, rtol=1e-12)
    assert np.all(s[~inner, 0] == 1.0) and np.all(b[~inner] == 0.0) and np.all(u[~inner] == 0.0)


def test_euler_axisymmetric_sources_on_uniform_state():
    """Kernel_TVD_old's 1/y terms (kernel.cu:3651-3662): on a uniform state
    one step is the forward-Euler step of the cylindrical Euler equations'
    geometric source -G/r, G = (rho v, rho u v, rho v^2, v (E + p))."""
    r, p, uu, vv = 1.2, 0.8, 0.35, -0.45
    gam = 1.4
    nx, ny, dx, dy, y0 = 7, 6, 0.1, 0.1, 0.35
    g = O.make_grid(nx, ny, dx=dx, dy=dy, y0=y0, gamma=gam, axisym=True)
    s = np.empty((ny, nx, 2)); u = np.empty((ny, nx, 2))
    s[...] = (r, p); u[...] = (uu, vv)
    dt = 2e-3
    for method in (0, 1, 2):
        s2, u2, _ = O.ref_euler_step("kernel_tvd_old", g, s, u, 1e7, dt, method)
        E = p / (gam - 1) + 0.5 * r * (uu * uu + vv * vv)
        for j in range(ny):
            y = y0 + j * dy
            G = np.array([r * vv, r * uu * vv, r * vv * vv, vv * (E + p)])
            W = np.array([r, r * uu, r * vv, E]) - dt * G / y
            ux, uy = W[1] / W[0], W[2] / W[0]
            pe = (W[3] - 0.5 * W[0] * (ux * ux + uy * uy)) * (gam - 1)
            want = np.array([W[0], pe, ux, uy])
            for i in range(nx):
                got = np.array([*s2[j, i], *u2[j, i]])
                assert np.allclose(got, want, rtol=1e-12, atol=1e-14), (method, j, i, got - want)


def test_euler_sphere_and_option_scope():
    nx, ny = 30, 20
    g = O.make_grid(nx, ny, dx=0.05, dy=0.05, gamma=1.4, r_frozen=0.42, axisym=True)
    rng = np.random.default_rng(4)
    s = np.empty((ny, nx, 2)); u = np.empty((ny, nx, 2))
    s[..., 0] = rng.uniform(0.5, 1.5, (ny, nx)); s[..., 1] = rng.uniform(0.5, 1.5, (ny, nx))
    u[...] = rng.uniform(-0.5, 0.5, (ny, nx, 2))
    x = g.x0 + np.arange(nx) * g.dx
    y = g.y0 + np.arange(ny) * g.dy
    X, Y = np.meshgrid(x, y)
    d = np.sqrt(X * X + Y * Y)
    assert np.min(np.abs(d - 0.42)) > 1e-6
    inside = d < 0.42
    s2, u2, _ = O.ref_euler_step("kernel_tvd_old", g, s, u, 1e7, 1e-3, 1)
    g.r_frozen = 0.0
    s2n, u2n, _ = O.ref_euler_step("kernel_tvd_old", g, s, u, 1e7, 1e-3, 1)
    np.testing.assert_array_equal(s2[inside], s[inside])
    np.testing.assert_array_equal(s2[~inside], s2n[~inside])
    np.testing.assert_array_equal(u2[~inside], u2n[~inside])
    # Ker_Dekard has neither option (kernel.cu:1736-1874)
    with pytest.raises(ValueError):
        O.ref_euler_step("ker_dekard", g, s, u, 1e7, 1e-3, 1)


def _mirror_rows(g, s, u, b=None):
    """The y-mirror image of a grid and its state: rows reversed, vy, By, Bz
    negated, the grid reflected about y = 0."""
    gm = O.make_grid(g.nx, g.ny, dx=g.dx, dy=g.dy, x0=g.x0, y0=-(g.y0 + (g.ny - 1) * g.dy), gamma=g.gamma,
                     axisym=bool(g.axisym))
    sm = np.ascontiguousarray(s[::-1])
    um = np.ascontiguousarray(u[::-1] * (1.0, -1.0))
    if b is None:
        return gm, sm, um
    return gm, sm, um, np.ascontiguousarray(b[::-1] * (1.0, -1.0, -1.0))


@pytest.mark.parametrize("method", [2, 3])
def test_add2_rows_below_axis_mirror_rows_above(method):
    h = 1.0 / 64.0
    g = O.make_grid(11, 9, dx=h, dy=h, x0=h / 2, y0=h / 2 + 3 * h, gamma=G53, axisym=True)
    s, u, b = smooth_state(g, seed=31 + method)
    gm, sm, um, bm = _mirror_rows(g, s, u, b)
    assert gm.y0 + (gm.ny - 1) * gm.dy < 0.0      # every row of the image is below the axis
    s2, u2, b2, T = O.ref_mhd_step("add2", g, s, u, b, 1e7, 2e-4, 0, method)
    s2m, u2m, b2m, Tm = O.ref_mhd_step("add2", gm, sm, um, bm, 1e7, 2e-4, 0, method)
    np.testing.assert_array_equal(s2m, s2[::-1])
    np.testing.assert_array_equal(u2m, u2[::-1] * (1.0, -1.0))
    np.testing.assert_array_equal(b2m, b2[::-1] * (1.0, -1.0, -1.0))
    assert Tm == T
    # the 1/y sources act on both grids
    g.axisym = 0
    s2c, _, _, _ = O.ref_mhd_step("add2", g, s, u, b, 1e7, 2e-4, 0, method)
    assert np.max(np.abs(s2c - s2)) > 1e-6


@pytest.mark.parametrize("method", [2, 3])
def test_add2_grid_across_the_axis_is_mirror_symmetric(method):
    """A state mirror-symmetric about y = 0 on a grid across the axis stays
    so: the rows below are the mirrored evaluation of the rows above."""
    h = 1.0 / 64.0
    nx, half = 10, 6
    g = O.make_grid(nx, 2 * half, dx=h, dy=h, x0=h / 2, y0=-(half - 0.5) * h, gamma=G53, axisym=True)
    gu = O.make_grid(nx, half, dx=h, dy=h, x0=h / 2, y0=h / 2, gamma=G53)
    s_up, u_up, b_up = smooth_state(gu, seed=5 + method)
    s = np.ascontiguousarray(np.concatenate([s_up[::-1], s_up]))
    u = np.ascontiguousarray(np.concatenate([u_up[::-1] * (1.0, -1.0), u_up]))
    b = np.ascontiguousarray(np.concatenate([b_up[::-1] * (1.0, -1.0, -1.0), b_up]))
    s2, u2, b2, _ = O.ref_mhd_step("add2", g, s, u, b, 1e7, 2e-4, 0, method)
    np.testing.assert_array_equal(s2[:half], s2[half:][::-1])
    np.testing.assert_array_equal(u2[:half], u2[half:][::-1] * (1.0, -1.0))
    np.testing.assert_array_equal(b2[:half], b2[half:][::-1] * (1.0, -1.0, -1.0))


@pytest.mark.parametrize("method", [2, 3])
def test_kernel_tvd_signed_y_mirror_rows_above(method):
    h = 1.0 / 64.0
    g = O.make_grid(12, 10, dx=h, dy=h, x0=h / 2, y0=h / 2 + 2 * h, gamma=G53, axisym=True)
    s, u, b = smooth_state(g, seed=17 + method)
    gm, sm, um, bm = _mirror_rows(g, s, u, b)
    s2, u2, b2, T = O.ref_mhd_step("kernel_tvd", g, s, u, b, 1e7, 2e-4, 0, method)
    s2m, u2m, b2m, Tm = O.ref_mhd_step("kernel_tvd", gm, sm, um, bm, 1e7, 2e-4, 0, method)
    assert np.max(np.abs(s2m - s2[::-1])) <= 1e-13
    assert np.max(np.abs(u2m - u2[::-1] * (1.0, -1.0))) <= 1e-13
    assert np.max(np.abs(b2m - b2[::-1] * (1.0, -1.0, -1.0))) <= 1e-13
    assert abs(Tm - T) <= 1e-14 * T


def test_kernel_tvd_old_signed_y_mirror_rows_above():
    h = 1.0 / 64.0
    g = O.make_grid(12, 10, dx=h, dy=h, x0=h / 2, y0=h / 2 + 2 * h, gamma=1.4, axisym=True)
    s, u, _ = smooth_state(g, seed=3)
    gm, sm, um = _mirror_rows(g, s, u)
    s2, u2, T = O.ref_euler_step("kernel_tvd_old", g, s, u, 1e7, 2e-4, 1)
    s2m, u2m, Tm = O.ref_euler_step("kernel_tvd_old", gm, sm, um, 1e7, 2e-4, 1)
    assert np.max(np.abs(s2m - s2[::-1])) <= 1e-13
    assert np.max(np.abs(u2m - u2[::-1] * (1.0, -1.0))) <= 1e-13
    assert abs(Tm - T) <= 1e-14 * T


def test_axisymmetric_row_on_the_axis_is_rejected():
    h = 0.125
    g = O.make_grid(6, 5, dx=h, dy=h, y0=-2 * h, gamma=G53, axisym=True)   # row 2 at y = 0
    s, u, b = smooth_state(g, seed=1)
    with pytest.raises(ValueError):
        O.ref_mhd_step("add2", g, s, u, b, 1e7, 1e-4, 0, 3)
    with pytest.raises(ValueError):
        O.ref_euler_step("kernel_tvd_old", g, s, u, 1e7, 1e-4, 1)


def _west_inflow_state(ny, nx, seed):
    """Rows whose west edge cell is subsonic with vx > -5 (the rule acts),
    supersonic (it does not), or subsonic with vx <= -5 (it does not)."""
    rng = np.random.default_rng(seed)
    s = np.stack([rng.uniform(0.6, 1.4, (ny, nx)), rng.uniform(0.6, 1.4, (ny, nx))], -1)
    u = rng.uniform(-0.4, 0.4, (ny, nx, 2))
    u[1, 0, 0] = 3.0                    # Mach > 1
    u[2, :2, 0] = -6.0                  # vx <= -5 ...
    s[2, :2, 1] = 80.0                  # ... and subsonic
    s[3, 0] = (1.0, 0.6)                # Mach 1.42 at its own pressure, 0.91 at
    u[3, 0] = (1.3, 0.0)                # 1/(gamma M_inf^2) with M_inf = 0.7, gamma = 1.4
    return np.ascontiguousarray(s), np.ascontiguousarray(u)


def _west_ghost(q_s, q_u, gamma, m_inf=None):
    """The ghost state of Kernel_TVD_old's west side from the edge cells
    (kernel.cu:3323-3337): vx = -5 where subsonic with vx > -5; with m_inf
    the pressure 1/(gamma m_inf^2), else the edge cell's own (kernel.cu:3365-3378)."""
    s, u = q_s.copy(), q_u.copy()
    mach = np.sqrt((u[:, 0] ** 2 + u[:, 1] ** 2) / (gamma * s[:, 1] / s[:, 0]))
    push = (u[:, 0] > -5.0) & (mach < 1)
    u[push, 0] = -5.0
    if m_inf is not None:
        s[:, 1] = 1.0 / (gamma * m_inf * m_inf)
    return s, u, push


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("axisym", [False, True])
def test_kernel_tvd_old_west_inflow_equals_explicit_ghost_columns(method, axisym):
    """Kernel_TVD_old's west ghost: column 0 and the columns from 2 on are the
    interior of the step on a grid two columns wider whose extra columns hold
    the ghost state; column 1 is that of a wider grid whose extra columns hold
    the edge cell under the vx rule with its own pressure; bit for bit."""
    nx, ny, h, gam, minf = 9, 7, 0.125, 1.4, 0.7
    x0, y0 = 0.0625, 0.0625
    bc = (O.BC_WEST_INFLOW, O.BC_FROZEN, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE)
    g = O.make_grid(nx, ny, dx=h, dy=h, x0=x0, y0=y0, gamma=gam, bc=bc, axisym=axisym, m_inf=minf)
    s, u = _west_inflow_state(ny, nx, seed=method)
    dt = 1e-3
    s2, u2, _ = O.ref_euler_step("kernel_tvd_old", g, s, u, 1e7, dt, method)
    wide_bc = (O.BC_TRANSMISSIVE, O.BC_FROZEN, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE)
    gw = O.make_grid(nx + 2, ny, dx=h, dy=h, x0=x0 - 2 * h, y0=y0, gamma=gam, bc=wide_bc, axisym=axisym)
    outs = []
    for m_inf in (minf, None):
        gs, gu, push = _west_ghost(s[:, 0], u[:, 0], gam, m_inf)
        S = np.ascontiguousarray(np.concatenate([gs[:, None], gs[:, None], s], axis=1))
        U = np.ascontiguousarray(np.concatenate([gu[:, None], gu[:, None], u], axis=1))
        S2, U2, _ = O.ref_euler_step("kernel_tvd_old", gw, S, U, 1e7, dt, method)
        outs.append((S2[:, 2:], U2[:, 2:]))
    assert push.any() and not push.all()
    (Sa, Ua), (Sb, Ub) = outs
    cols_a = [0] + list(range(2, nx))
    np.testing.assert_array_equal(s2[:, cols_a], Sa[:, cols_a])
    np.testing.assert_array_equal(u2[:, cols_a], Ua[:, cols_a])
    np.testing.assert_array_equal(s2[:, 1], Sb[:, 1])
    np.testing.assert_array_equal(u2[:, 1], Ub[:, 1])
    # the two views of column -1 differ, and the side differs from a copy
    assert np.any(Sa[:, 1] != Sb[:, 1])
    gt = O.make_grid(nx, ny, dx=h, dy=h, x0=x0, y0=y0, gamma=gam, bc=wide_bc, axisym=axisym)
    s2t, _, _ = O.ref_euler_step("kernel_tvd_old", gt, s, u, 1e7, dt, method)
    assert np.any(s2t[:, 0] != s2[:, 0]) and np.any(s2t[:, 1] != s2[:, 1])


def test_west_inflow_scope():
    """Kernel_TVD_old only, west side only, M_inf > 0."""
    nx, ny = 6, 5
    s, u = _west_inflow_state(ny, nx, seed=0)
    b = np.zeros((ny, nx, 3))
    ok = (O.BC_WEST_INFLOW, O.BC_FROZEN, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE)
    O.ref_euler_step("kernel_tvd_old", O.make_grid(nx, ny, bc=ok, m_inf=0.5), s, u, 1e7, 1e-4, 1)
    for bad in (dict(bc=ok, m_inf=0.0),
                dict(bc=(O.BC_TRANSMISSIVE, O.BC_WEST_INFLOW, O.BC_TRANSMISSIVE, O.BC_TRANSMISSIVE), m_inf=0.5)):
        with pytest.raises(ValueError):
            O.ref_euler_step("kernel_tvd_old", O.make_grid(nx, ny, **bad), s, u, 1e7, 1e-4, 1)
    with pytest.raises(ValueError):
        O.ref_euler_step("ker_dekard", O.make_grid(nx, ny, bc=ok, m_inf=0.5), s, u, 1e7, 1e-4, 1)
    with pytest.raises(ValueError):
        O.ref_mhd_step("add2", O.make_grid(nx, ny, bc=ok, m_inf=0.5, gamma=G53), s, u, b, 1e7, 1e-4, 0, 3)
