"""Pin the oracle's assembled steps (stencil, boundary kinds, update, dt).

Checks: free-stream preservation, conservation over a periodic box, exact
y-invariance of x-directed data, periodic translation invariance, reflective
mirror symmetry, L1 convergence of first-order Godunov to the exact Sod
solution, agreement of the reference-layout (AoS primitive) steps with the
native (SoA conserved) stage, the SSP-RK2 combination, and the CFL dt against
a brute-force numpy evaluation of the cell-centred wave-speed bound.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import textbook as TB
from golikov_2d_amd import problems as PB


def euler_scheme(flux=O.FLUX_HLLC, order=2, cfl=0.4):
    return O.Scheme(O.PHYS_EULER, flux, order, cfl)


def mhd_scheme(flux=O.FLUX_HLLD, order=2, cfl=0.3):
    return O.Scheme(O.PHYS_MHD, flux, order, cfl)


@pytest.mark.parametrize("flux,order", [(O.FLUX_HLLC, 1), (O.FLUX_HLLC, 2), (O.FLUX_EXACT, 1), (O.FLUX_EXACT, 2)])
def test_free_stream_preserved(flux, order):
    g = O.make_grid(12, 9, bc=(O.BC_TRANSMISSIVE, O.BC_PERIODIC, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE))
    # reflective sides need zero normal velocity for a free stream
    g.bc[2] = O.BC_TRANSMISSIVE
    U = PB.uniform_euler(g, rho=0.7, p=1.3, u=0.4, v=-0.25)
    U1 = O.stage(g, euler_scheme(flux, order), U, None, 0.0, 1.0, 1e-3)
    assert np.array_equal(U1, U) or np.max(np.abs(U1 - U)) < 1e-15


def test_mhd_free_stream_preserved():
    g = O.make_grid(10, 8, bc=(O.BC_PERIODIC,) * 4)
    U = PB.uniform_mhd(g, rho=1.0, p=0.6, u=0.3, v=-0.2, B=(0.5, -0.3, 0.2))
    for flux in (O.FLUX_HLLD, O.FLUX_HLLC_MHD):
        U1 = O.stage(g, mhd_scheme(flux), U, None, 0.0, 1.0, 1e-3)
        assert np.max(np.abs(U1 - U)) < 1e-14


@pytest.mark.parametrize("flux,order", [(O.FLUX_HLLC, 2), (O.FLUX_EXACT, 1)])
def test_conservation_periodic(flux, order):
    g = O.make_grid(32, 24, bc=(O.BC_PERIODIC,) * 4)
    U = PB.four_quadrant(g, seed=3)
    tot0 = U.sum(axis=(1, 2))
    steps, _ = O.run(g, euler_scheme(flux, order), 2, U, nsteps=5)
    assert steps == 5
    tot1 = U.sum(axis=(1, 2))
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * np.abs(U).sum(axis=(1, 2)))


def test_mhd_conservation_periodic_mass_energy():
    # the Powell source acts on momentum, B and E; mass is conserved exactly
    g = O.make_grid(32, 32, bc=(O.BC_PERIODIC,) * 4, gamma=5 / 3)
    U = PB.orszag_tang(g, seed=0, amp=0.0)
    m0 = U[0].sum()
    O.run(g, mhd_scheme(), 2, U, nsteps=3)
    assert abs(U[0].sum() - m0) <= 1e-13 * m0


def test_x_directed_problem_stays_y_invariant():
    g = O.make_grid(64, 6, bc=(O.BC_TRANSMISSIVE,) * 4)
    U = PB.sod_x(g)
    O.run(g, euler_scheme(O.FLUX_HLLC, 2), 2, U, nsteps=20)
    assert np.all(U == U[:, :1, :])
    assert np.all(U[2] == 0.0)


def test_periodic_translation_invariance():
    g = O.make_grid(24, 20, bc=(O.BC_PERIODIC,) * 4)
    U = PB.four_quadrant(g, seed=7)
    Us = np.roll(U, (3, 5), axis=(1, 2)).copy()
    sc = euler_scheme()
    A = O.stage(g, sc, U, None, 0.0, 1.0, 2e-3)
    B = O.stage(g, sc, Us, None, 0.0, 1.0, 2e-3)
    assert np.array_equal(np.roll(A, (3, 5), axis=(1, 2)), B)


def test_reflective_mirror_symmetry():
    # data symmetric about the west wall mirrored: x -> -x, u -> -u
    g = O.make_grid(16, 8, bc=(O.BC_REFLECTIVE, O.BC_TRANSMISSIVE, O.BC_TRANSMISSIVE, O.BC_TRANSMISSIVE))
    gm = O.make_grid(32, 8, x0=-1.0 + 0.5 / 16, dx=1.0 / 16,
                     bc=(O.BC_TRANSMISSIVE,) * 4)
    U = PB.four_quadrant(g, seed=1)
    # full domain = mirror image on the left + original on the right
    Um = np.concatenate([U[:, :, ::-1], U], axis=2).copy()
    Um[1, :, :16] *= -1.0
    sc = euler_scheme(O.FLUX_HLLC, 1)
    A = O.stage(g, sc, U, None, 0.0, 1.0, 1e-3)
    B = O.stage(gm, sc, Um, None, 0.0, 1.0, 1e-3)
    assert np.max(np.abs(A - B[:, :, 16:])) < 1e-14


def test_sod_godunov_converges_to_exact_solution():
    errs = []
    for nx in (100, 200, 400):
        g = O.make_grid(nx, 2, gamma=1.4)
        U = PB.sod_x(g)
        steps, t = O.run(g, euler_scheme(O.FLUX_EXACT, 1), 1, U, t_end=0.2)
        assert abs(t - 0.2) < 1e-15
        x = (np.arange(nx) + 0.5) / nx
        ex = np.array([r for r, _, _ in TB.sod_exact(1.4, x, 0.2)])
        errs.append(np.mean(np.abs(U[0, 0] - ex)))
    # first order on discontinuities: error ~ h^(1/2..1)
    assert errs[0] < 0.02 and errs[1] < errs[0] * 0.8 and errs[2] < errs[1] * 0.8, errs


def test_sod_muscl_hllc_more_accurate_than_first_order():
    nx = 200
    x = (np.arange(nx) + 0.5) / nx
    ex = np.array([r for r, _, _ in TB.sod_exact(1.4, x, 0.2)])
    e = {}
    for order in (1, 2):
        g = O.make_grid(nx, 2)
        U = PB.sod_x(g)
        O.run(g, euler_scheme(O.FLUX_HLLC, order), 2, U, t_end=0.2)
        e[order] = np.mean(np.abs(U[0, 0] - ex))
    assert e[2] < 0.7 * e[1], e


@pytest.mark.parametrize("kind,order", [("ker_dekard", 1), ("kernel_tvd_old", 2)])
@pytest.mark.parametrize("method", [1, 2])
def test_reference_layout_matches_native_stage(kind, order, method):
    g = O.make_grid(20, 16, bc=(O.BC_TRANSMISSIVE, O.BC_PERIODIC, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE))
    U = PB.four_quadrant(g, seed=5)
    s, u, _ = O.cons_to_prim(g, O.PHYS_EULER, U)
    dt = 2e-4
    s2, u2, T = O.ref_euler_step(kind, g, s, u, 1e7, dt, method)
    flux = O.FLUX_EXACT if method == 2 else O.FLUX_HLLC
    U1 = O.stage(g, euler_scheme(flux, order), U, None, 0.0, 1.0, dt)
    s1, u1, _ = O.cons_to_prim(g, O.PHYS_EULER, U1)
    assert np.max(np.abs(s2 - s1) / np.abs(s1)) < 1e-12
    assert np.max(np.abs(u2 - u1)) < 1e-12
    assert 0 < T < 1e7


@pytest.mark.parametrize("kind,order", [("add2", 1), ("kernel_tvd", 2)])
@pytest.mark.parametrize("method", [2, 3])
def test_reference_mhd_layout_matches_native_stage(kind, order, method):
    g = O.make_grid(16, 16, bc=(O.BC_PERIODIC,) * 4, gamma=5 / 3, krit=0.25)
    U = PB.orszag_tang(g, seed=2, amp=1e-3)
    s, u, b = O.cons_to_prim(g, O.PHYS_MHD, U)
    dt = 1e-4
    s2, u2, b2, T = O.ref_mhd_step(kind, g, s, u, b, 1e7, dt, 0, method)
    flux = O.FLUX_HLLD if method == 3 else O.FLUX_HLLC_MHD
    U1 = O.stage(g, mhd_scheme(flux, order), U, None, 0.0, 1.0, dt)
    s1, u1, b1 = O.cons_to_prim(g, O.PHYS_MHD, U1)
    assert np.max(np.abs(s2 - s1) / np.abs(s1)) < 1e-11
    assert np.max(np.abs(u2 - u1)) < 1e-11
    assert np.max(np.abs(b2 - b1)) < 1e-12
    assert 0 < T < 1e7


def test_ssprk2_is_average_of_base_and_two_euler_stages():
    g = O.make_grid(18, 14, bc=(O.BC_PERIODIC,) * 4)
    U0 = PB.four_quadrant(g, seed=9)
    sc = euler_scheme()
    dt = 1e-3
    U1 = O.stage(g, sc, U0, None, 0.0, 1.0, dt)
    U2 = O.stage(g, sc, U1, None, 0.0, 1.0, dt)
    R = O.stage(g, sc, U1, U0, 0.5, 0.5, dt)
    assert np.max(np.abs(R - (0.5 * U0 + 0.5 * U2))) < 1e-15 * np.max(np.abs(U0)) * 4


def test_cfl_dt_matches_bruteforce():
    g = O.make_grid(40, 30, dx=1 / 40, dy=1 / 60)
    U = PB.four_quadrant(g, seed=4)
    r, mx, my, E = U
    u, v = mx / r, my / r
    p = 0.4 * (E - 0.5 * (mx * u + my * v))
    c = np.sqrt(1.4 * p / r)
    w = ((np.abs(u) + c) * 40 + (np.abs(v) + c) * 60).max()
    assert abs(O.wavespeed_max(g, euler_scheme(), U) - w) <= 1e-15 * w
    g2 = O.make_grid(40, 30, dx=1 / 40, dy=1 / 60, bc=(O.BC_PERIODIC,) * 4)
    _, wo = O.stage(g2, euler_scheme(), U, None, 0.0, 1.0, 1e-4, want_w=True)
    U1 = O.stage(g2, euler_scheme(), U, None, 0.0, 1.0, 1e-4)
    assert wo == O.wavespeed_max(g2, euler_scheme(), U1)


def test_funk_time_semantics():
    # kernel.cu:1722-1733
    T, T_do, TT, i = O.funk_time(0.01, 1e-8, 0.5, 4)
    assert (T, T_do, TT, i) == (1e7, 0.01, 0.51, 5)


def test_no_clamps_or_caps_on_quadrant_runs():
    O.diag_reset()
    g = O.make_grid(48, 40)
    U = PB.four_quadrant(g, seed=12)
    O.run(g, euler_scheme(), 2, U, nsteps=20)
    d = O.diag()
    assert all(v == 0 for v in d.values()), d
    assert np.all(U[0] > 0)


def test_exact_solver_run_has_no_iteration_cap_hits():
    O.diag_reset()
    g = O.make_grid(40, 40)
    U = PB.four_quadrant(g, seed=13)
    O.run(g, euler_scheme(O.FLUX_EXACT, 1), 1, U, nsteps=10)
    d = O.diag()
    assert d["newton_cap_hits"] == 0 and d["bisect_cap_hits"] == 0 and d["bad_branch"] == 0
