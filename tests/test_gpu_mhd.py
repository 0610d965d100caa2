"""GPU parity of the ideal-MHD path (HLLC-MHD / HLLD, Powell source) through
the C ABI against the CPU oracle.

Bar (per conserved field, relative L1): one stage from identical inputs
<= 1e-12; 100 SSP-RK2 steps <= 1e-10; positive density and pressure. At the
2048^2 Orszag-Tang size: one full stage against the oracle, and the point
symmetry (x, y) -> (1-x, 1-y), v -> -v, B -> -B of the unperturbed vortex kept
to 1e-12 over 20 steps.
"""
import numpy as np
import pytest
import torch

from golikov_2d_amd import api
from golikov_2d_amd import problems as PB
from oracle import oracle as O
from tests.gpu_util import dev, grids, rel_l1, say, to_dev, to_host

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12
RUN_TOL = 1e-10
G53 = 5.0 / 3.0

BCS = {
    "periodic": (O.BC_PERIODIC,) * 4,
    "transmissive": (O.BC_TRANSMISSIVE,) * 4,
    "mixed": (O.BC_REFLECTIVE, O.BC_TRANSMISSIVE, O.BC_PERIODIC, O.BC_PERIODIC),
}
FLUX = {3: (O.FLUX_HLLD, api.FLUX_HLLD), 2: (O.FLUX_HLLC_MHD, api.FLUX_HLLC_MHD)}


def oscheme(metod, order, cfl=0.3):
    return O.Scheme(O.PHYS_MHD, FLUX[metod][0], order, cfl)


def pscheme(metod, order, integ=api.SSPRK2, cfl=0.3):
    return api.make_scheme(api.MHD, FLUX[metod][1], order, integ, cfl)


@pytest.mark.parametrize("shape", [(130, 97), (64, 63), (1, 6)])
@pytest.mark.parametrize("bc", list(BCS))
@pytest.mark.parametrize("metod,order", [(3, 2), (3, 1), (2, 2), (2, 1)])
def test_mhd_stage_matches_oracle(shape, bc, metod, order):
    nx, ny = shape
    say(f"mhd stage {nx}x{ny} bc={bc} metod={metod} order={order}")
    g, og = grids(nx, ny, BCS[bc], gamma=G53)
    U = PB.orszag_tang(og, seed=nx + ny, amp=1e-2)
    dt = 0.05 / (nx + ny)
    ref, wref = O.stage(og, oscheme(metod, order), U, None, 0.0, 1.0, dt, want_w=True)
    Ud = to_dev(U)
    out = torch.empty_like(Ud)
    w = torch.zeros(1, dtype=torch.int64, device=dev())
    api.stage(g, pscheme(metod, order), Ud, None, out, 0.0, 1.0, dt, wmax_bits=w)
    got = to_host(out)
    err = rel_l1(got, ref)
    # Bz is identically zero for this initial state in both
    assert np.all(err <= STEP_TOL), err
    assert abs(api.bits_to_double(w) - wref) <= 1e-14 * wref


@pytest.mark.parametrize("metod", [3, 2])
def test_mhd_100_step_run_matches_oracle(metod):
    n = 96
    say(f"mhd 100-step run {n}^2 metod={metod}")
    g, og = grids(n, n, BCS["periodic"], gamma=G53)
    U = PB.orszag_tang(og, seed=4, amp=1e-3)
    Ud = to_dev(U)
    steps, t = O.run(og, oscheme(metod, 2), 2, U, nsteps=100)
    api.diagnostics(reset=True)
    s = api.Solver(g, pscheme(metod, 2))
    s.start(Ud)
    s.advance(Ud, 100)
    tg, ng, _ = s.time()
    got = to_host(Ud)
    assert ng == steps == 100 and abs(tg - t) <= 1e-12 * t
    err = rel_l1(got, U)
    assert np.all(err <= RUN_TOL), err
    d = api.diagnostics()
    assert d["clamp_rho"] == 0 and d["clamp_p"] == 0
    s.close()


@pytest.mark.parametrize("kind,order", [("add2", 1), ("kernel_tvd", 2)])
@pytest.mark.parametrize("method", [2, 3])
def test_reference_mhd_layout_matches_oracle(kind, order, method):
    say(f"reference layout {kind} method={method}")
    bc = (O.BC_TRANSMISSIVE, O.BC_FROZEN, O.BC_REFLECTIVE, O.BC_FROZEN)
    g, og = grids(131, 70, bc, gamma=G53, krit=0.25)
    U = PB.orszag_tang(og, seed=6, amp=1e-2)
    s, u, b = O.cons_to_prim(og, O.PHYS_MHD, U)
    T0, Tdo = 1e7, 2e-4
    s2r, u2r, b2r, Tr = O.ref_mhd_step(kind, og, s, u, b, T0, Tdo, 0, method)
    sd, ud, bd = to_dev(s), to_dev(u), to_dev(b)
    s2, u2, b2 = torch.empty_like(sd), torch.empty_like(ud), torch.empty_like(bd)
    T = torch.tensor([T0], dtype=torch.float64, device=dev())
    T_do = torch.tensor([Tdo], dtype=torch.float64, device=dev())
    fn = api.add2 if kind == "add2" else api.kernel_tvd
    fn(g, sd, ud, bd, s2, u2, b2, T, T_do, 0, method)
    assert np.max(np.abs(to_host(s2) - s2r) / np.abs(s2r)) <= 1e-12
    assert np.max(np.abs(to_host(u2) - u2r)) <= 1e-12
    assert np.max(np.abs(to_host(b2) - b2r)) <= 1e-12
    assert abs(T.item() - Tr) <= 1e-14 * Tr


def test_orszag_tang_full_size_stage_parity_2048():
    n = 2048
    say("mhd 2048^2 stage parity")
    g, og = grids(n, n, BCS["periodic"], gamma=G53)
    U = PB.orszag_tang(og, seed=0, amp=1e-3)
    dt = 0.3 / O.wavespeed_max(og, oscheme(3, 2), U)
    ref = O.stage(og, oscheme(3, 2), U, None, 0.0, 1.0, dt)
    Ud = to_dev(U)
    out = torch.empty_like(Ud)
    api.stage(g, pscheme(3, 2), Ud, None, out, 0.0, 1.0, dt)
    err = rel_l1(to_host(out), ref)
    assert np.all(err[:5] <= STEP_TOL) and err[6] <= STEP_TOL, err


def test_orszag_tang_point_symmetry_2048():
    n = 2048
    say("mhd 2048^2 point symmetry")
    g = api.make_grid(n, n, gamma=G53, bc=(api.BC_PERIODIC,) * 4)
    og = O.make_grid(n, n, gamma=G53, bc=(O.BC_PERIODIC,) * 4)
    U = to_dev(PB.orszag_tang(og, seed=0, amp=0.0))
    s = api.Solver(g, pscheme(3, 2))
    s.start(U)
    s.advance(U, 20)
    rot = torch.flip(U, dims=(1, 2))
    sign = torch.tensor([1, -1, -1, -1, -1, -1, 1], dtype=torch.float64, device=dev()).view(7, 1, 1)
    diff = (U - sign * rot).abs().sum(dim=(1, 2)) / U.abs().sum(dim=(1, 2)).clamp_min(1e-300)
    assert torch.all(diff[[0, 1, 2, 3, 4, 6]] <= 1e-12), diff
    s.close()


PROD_BC = (O.BC_SYMMETRY, O.BC_OUTFLOW, O.BC_AXIS, O.BC_INFLOW)


def _ref_step_dev(kind, g, s, u, b, T0, Tdo, method):
    sd, ud, bd = to_dev(s), to_dev(u), to_dev(b)
    s2, u2, b2 = torch.empty_like(sd), torch.empty_like(ud), torch.empty_like(bd)
    T = torch.tensor([T0], dtype=torch.float64, device=dev())
    T_do = torch.tensor([Tdo], dtype=torch.float64, device=dev())
    fn = api.add2 if kind == "add2" else api.kernel_tvd
    fn(g, sd, ud, bd, s2, u2, b2, T, T_do, 0, method)
    return to_host(s2), to_host(u2), to_host(b2), T.item()


@pytest.mark.parametrize("kind", ["add2", "kernel_tvd"])
@pytest.mark.parametrize("method", [2, 3])
@pytest.mark.parametrize("axisym", [True, False])
def test_production_options_match_oracle(kind, method, axisym):
    """Production side kinds, frozen inner sphere and (optionally) the
    axisymmetric sources: one step against the oracle, 1e-12."""
    from tests.test_oracle_wind import smooth_state
    nx, ny = 203, 131
    say(f"production options {kind} method={method} axisym={axisym} {nx}x{ny}")
    kw = PB.wind_grid_args(nx, ny, x_max=4.0, y_max=2.6)
    kw["axisym"] = axisym
    g, og = grids(nx, ny, kw.pop("bc"), **kw)
    s, u, b = smooth_state(og, seed=method + 7)
    T0, Tdo = 1e7, 1e-3
    s2r, u2r, b2r, Tr = O.ref_mhd_step(kind, og, s, u, b, T0, Tdo, 0, method)
    s2, u2, b2, T = _ref_step_dev(kind, g, s, u, b, T0, Tdo, method)
    assert np.max(np.abs(s2 - s2r) / np.abs(s2r)) <= 1e-12
    assert np.max(np.abs(u2 - u2r)) <= 1e-12
    assert np.max(np.abs(b2 - b2r)) <= 1e-12
    assert abs(T - Tr) <= 1e-14 * Tr
    # the frozen sphere is copied bit for bit
    x = og.x0 + np.arange(nx) * og.dx
    y = og.y0 + np.arange(ny) * og.dy
    X, Y = np.meshgrid(x, y)
    inside = np.sqrt(X * X + Y * Y) < og.r_frozen
    assert inside.sum() > 10
    np.testing.assert_array_equal(s2[inside], s[inside])


def test_production_options_rejected_elsewhere():
    """The production side kinds and options are accepted by the MHD
    reference-layout steps only, each side kind on its own side."""
    say("production options rejected elsewhere")
    kw = PB.wind_grid_args(64, 40)
    g = api.make_grid(64, 40, **kw)
    U = to_dev(PB.uniform_mhd(g))
    out = torch.empty_like(U)
    with pytest.raises(api.GD2Error):
        api.stage(g, pscheme(3, 2), U, None, out, 0.0, 1.0, 1e-4)
    sw = list(kw["bc"])
    sw[0], sw[1] = sw[1], sw[0]
    kw2 = dict(kw, bc=tuple(sw))
    g2 = api.make_grid(64, 40, **kw2)
    s = torch.ones((40, 64, 2), dtype=torch.float64, device=dev())
    u = torch.zeros((40, 64, 2), dtype=torch.float64, device=dev())
    b = torch.zeros((40, 64, 3), dtype=torch.float64, device=dev())
    T = torch.tensor([1e7], dtype=torch.float64, device=dev())
    with pytest.raises(api.GD2Error):
        api.kernel_tvd(g2, s, u, b, torch.empty_like(s), torch.empty_like(u), torch.empty_like(b), T, T.clone(), 0, 2)


def test_stellar_wind_20_steps_match_oracle():
    """The production run at a tenth of its size: stellar-wind state, add2
    (first-order) then Kernel_TVD steps, each followed by funk_time, as the
    reference's step loop does; 20 steps against the oracle, 1e-10."""
    nx, ny = 322, 179
    say(f"stellar wind {nx}x{ny} 20 steps")
    kw = PB.wind_grid_args(nx, ny)
    g, og = grids(nx, ny, kw.pop("bc"), **kw)
    s, u, b = PB.stellar_wind_prim(og)
    # host loop through the oracle
    hs, hu, hb = s.copy(), u.copy(), b.copy()
    T, Tdo, TT, it = 1e7, 1e-8, 0.0, 0
    for k in range(20):
        kind = "add2" if k < 10 else "kernel_tvd"
        hs, hu, hb, T = O.ref_mhd_step(kind, og, hs, hu, hb, T, Tdo, it, 2)
        T, Tdo, TT, it = O.funk_time(T, Tdo, TT, it)
    # device loop through the C ABI
    sd, ud, bd = to_dev(s), to_dev(u), to_dev(b)
    s2, u2, b2 = torch.empty_like(sd), torch.empty_like(ud), torch.empty_like(bd)
    Td = torch.tensor([1e7], dtype=torch.float64, device=dev())
    Tdod = torch.tensor([1e-8], dtype=torch.float64, device=dev())
    TTd = torch.zeros(1, dtype=torch.float64, device=dev())
    itd = torch.zeros(1, dtype=torch.int32, device=dev())
    for k in range(20):
        fn = api.add2 if k < 10 else api.kernel_tvd
        fn(g, sd, ud, bd, s2, u2, b2, Td, Tdod, 0, 2)
        sd, s2, ud, u2, bd, b2 = s2, sd, u2, ud, b2, bd
        api.funk_time(Td, Tdod, TTd, itd)
    gs, gu, gb = to_host(sd), to_host(ud), to_host(bd)
    assert abs(TTd.item() - TT) <= 1e-12 * TT and itd.item() == it
    assert np.max(np.abs(gs - hs) / np.abs(hs)) <= 1e-10
    assert np.max(np.abs(gu - hu)) <= 1e-10
    assert np.max(np.abs(gb - hb)) <= 1e-10
    assert np.all(gs[..., 1] > 0)


def test_stellar_wind_full_size_step_matches_oracle():
    """The production run at its own size (3226 x 1792): one Kernel_TVD
    (HLLC-MHD, method 2) step from the stellar-wind state after 30 add2
    steps, against the oracle, 1e-12; the frozen sphere copied bit for bit."""
    nx, ny = PB.WIND_NX, PB.WIND_NY
    say(f"stellar wind full size {nx}x{ny}")
    kw = PB.wind_grid_args()
    g, og = grids(nx, ny, kw.pop("bc"), **kw)
    s, u, b = PB.stellar_wind_prim(og)
    sd, ud, bd = to_dev(s), to_dev(u), to_dev(b)
    s2, u2, b2 = torch.empty_like(sd), torch.empty_like(ud), torch.empty_like(bd)
    Td = torch.tensor([1e7], dtype=torch.float64, device=dev())
    Tdod = torch.tensor([1e-8], dtype=torch.float64, device=dev())
    TTd = torch.zeros(1, dtype=torch.float64, device=dev())
    itd = torch.zeros(1, dtype=torch.int32, device=dev())
    for _ in range(30):
        api.add2(g, sd, ud, bd, s2, u2, b2, Td, Tdod, 0, 2)
        sd, s2, ud, u2, bd, b2 = s2, sd, u2, ud, b2, bd
        api.funk_time(Td, Tdod, TTd, itd)
    s0, u0, b0 = to_host(sd), to_host(ud), to_host(bd)
    T0, Tdo = Td.item(), Tdod.item()
    s2r, u2r, b2r, Tr = O.ref_mhd_step("kernel_tvd", og, s0, u0, b0, T0, Tdo, 0, 2)
    api.kernel_tvd(g, sd, ud, bd, s2, u2, b2, Td, Tdod, 0, 2)
    gs, gu, gb = to_host(s2), to_host(u2), to_host(b2)
    assert np.max(np.abs(gs - s2r) / np.abs(s2r)) <= 1e-12
    assert np.max(np.abs(gu - u2r)) <= 1e-12
    assert np.max(np.abs(gb - b2r)) <= 1e-12
    assert abs(Td.item() - Tr) <= 1e-14 * Tr
    x = og.x0 + np.arange(nx) * og.dx
    y = og.y0 + np.arange(ny) * og.dy
    X, Y = np.meshgrid(x, y)
    inside = np.sqrt(X * X + Y * Y) < og.r_frozen
    np.testing.assert_array_equal(gs[inside], s0[inside])


def test_reference_layout_16384_square_rows_equal_one_row():
    """64-bit indexing of the AoS reference layout at 16384^2 (the
    reference's int byte counts overflow above ~1.3e8 cells; b alone is
    6.4 GB here): an x-directed MHD state, constant along y with transmissive
    sides, gives every row of a Kernel_TVD (HLLD) step bit for bit equal to
    the step on 16384 x 4, and the same time step; that row matches the
    oracle to 1e-12."""
    n = 16384
    say("reference layout 16384^2 rows")
    x = (np.arange(n) + 0.5) / n
    s_row = np.stack([1.0 + 0.3 * np.sin(6 * np.pi * x) + np.where(x < 0.5, 0.5, 0.0),
                      0.8 + 0.2 * np.cos(4 * np.pi * x)], axis=-1)
    u_row = np.stack([0.3 * np.sin(2 * np.pi * x), np.full(n, 0.1)], axis=-1)
    b_row = np.stack([np.full(n, 0.7), 0.4 * np.cos(2 * np.pi * x), 0.3 * np.sin(4 * np.pi * x)], axis=-1)
    res = {}
    for ny in (4, n):
        g = api.make_grid(n, ny, dx=1.0 / n, dy=1.0 / n, gamma=G53, krit=0.25)
        sd = to_dev(s_row[None]).repeat(ny, 1, 1).contiguous()
        ud = to_dev(u_row[None]).repeat(ny, 1, 1).contiguous()
        bd = to_dev(b_row[None]).repeat(ny, 1, 1).contiguous()
        s2, u2, b2 = torch.empty_like(sd), torch.empty_like(ud), torch.empty_like(bd)
        T = torch.tensor([1e7], dtype=torch.float64, device=dev())
        T_do = torch.tensor([2e-5], dtype=torch.float64, device=dev())
        api.kernel_tvd(g, sd, ud, bd, s2, u2, b2, T, T_do, 0, 3)
        torch.cuda.synchronize()
        del sd, ud, bd
        res[ny] = (s2, u2, b2, T.item())
    s4, u4, b4, T4 = res[4]
    sn, un, bn, Tn = res[n]
    assert T4 == Tn
    for big, small in ((sn, s4), (un, u4), (bn, b4)):
        assert torch.equal(big, small[:1].expand_as(big))
    og = O.make_grid(n, 4, dx=1.0 / n, dy=1.0 / n, gamma=G53, krit=0.25)
    rep = lambda a: np.ascontiguousarray(np.repeat(a[None], 4, axis=0))
    s2r, u2r, b2r, Tr = O.ref_mhd_step("kernel_tvd", og, rep(s_row), rep(u_row), rep(b_row), 1e7, 2e-5, 0, 3)
    assert np.max(np.abs(to_host(s4) - s2r) / np.abs(s2r)) <= 1e-12
    assert np.max(np.abs(to_host(b4) - b2r)) <= 1e-12
    assert abs(T4 - Tr) <= 1e-14 * Tr


@pytest.mark.parametrize("shape", [(1, 5), (5, 1), (2, 2), (3, 4), (66, 3)])
@pytest.mark.parametrize("kind", ["add2", "kernel_tvd"])
def test_production_options_small_grids_match_oracle(shape, kind):
    """Production sides on grids one or a few cells wide (ghost layers that
    mirror across the whole grid, edge cells that are ghosts on two sides),
    with the axisymmetric sources: one step against the oracle, 1e-12."""
    from tests.test_oracle_wind import smooth_state
    nx, ny = shape
    say(f"production options small grid {kind} {nx}x{ny}")
    g, og = grids(nx, ny, (O.BC_SYMMETRY, O.BC_OUTFLOW, O.BC_AXIS, O.BC_INFLOW), gamma=G53, dx=0.05, dy=0.05,
                  x0=0.025, y0=0.025, axisym=True)
    s, u, b = smooth_state(og, seed=nx * 7 + ny)
    T0, Tdo = 1e7, 1e-4
    for method in (2, 3):
        s2r, u2r, b2r, Tr = O.ref_mhd_step(kind, og, s, u, b, T0, Tdo, 0, method)
        s2, u2, b2, T = _ref_step_dev(kind, g, s, u, b, T0, Tdo, method)
        assert np.max(np.abs(s2 - s2r) / np.abs(s2r)) <= 1e-12, method
        assert np.max(np.abs(u2 - u2r)) <= 1e-12, method
        assert np.max(np.abs(b2 - b2r)) <= 1e-12, method
        assert abs(T - Tr) <= 1e-14 * Tr, method


@pytest.mark.parametrize("kind", ["add2", "kernel_tvd"])
@pytest.mark.parametrize("method", [2, 3])
def test_rows_below_the_axis_match_oracle(kind, method):
    """Axisymmetric steps on a grid across the axis (rows at y < 0): add2's
    mirrored evaluation of those rows (kernel.cu:1885-1889, 2026-2057,
    2177-2182) and Kernel_TVD's signed 1/y terms, one step against the
    oracle, 1e-12; a row centred on the axis is rejected."""
    from tests.test_oracle_wind import smooth_state
    nx, ny, h = 131, 70, 0.05
    say(f"rows below the axis {kind} method={method} {nx}x{ny}")
    g, og = grids(nx, ny, (O.BC_TRANSMISSIVE,) * 4, gamma=G53, dx=h, dy=h, x0=h / 2, y0=-29.5 * h, axisym=True)
    s, u, b = smooth_state(og, seed=method + 41)
    T0, Tdo = 1e7, 2e-4
    s2r, u2r, b2r, Tr = O.ref_mhd_step(kind, og, s, u, b, T0, Tdo, 0, method)
    s2, u2, b2, T = _ref_step_dev(kind, g, s, u, b, T0, Tdo, method)
    assert np.max(np.abs(s2 - s2r) / np.abs(s2r)) <= 1e-12
    assert np.max(np.abs(u2 - u2r)) <= 1e-12
    assert np.max(np.abs(b2 - b2r)) <= 1e-12
    assert abs(T - Tr) <= 1e-14 * Tr
    g0 = api.make_grid(nx, ny, gamma=G53, dx=h, dy=h, x0=h / 2, y0=-30 * h, axisym=True)
    with pytest.raises(api.GD2Error):
        _ref_step_dev(kind, g0, s, u, b, T0, Tdo, method)


@pytest.mark.parametrize("method", [2, 3])
def test_rows_below_the_axis_20_steps_match_oracle(method):
    """add2 then Kernel_TVD steps with funk_time, as the reference's loop
    drives them, on a grid across the axis (the oracle evaluates add2's rows
    below it in the mirrored frame, the kernels with the signed y): 20 steps
    against the oracle, 1e-10."""
    from tests.test_oracle_wind import smooth_state
    nx, ny, h = 131, 70, 0.05
    say(f"rows below the axis 20 steps method={method}")
    g, og = grids(nx, ny, (O.BC_TRANSMISSIVE,) * 4, gamma=G53, dx=h, dy=h, x0=h / 2, y0=-29.5 * h, axisym=True)
    s, u, b = smooth_state(og, seed=9)
    hs, hu, hb = s.copy(), u.copy(), b.copy()
    T, Tdo, TT, it = 1e7, 1e-8, 0.0, 0
    for k in range(20):
        kind = "add2" if k < 10 else "kernel_tvd"
        hs, hu, hb, T = O.ref_mhd_step(kind, og, hs, hu, hb, T, Tdo, it, method)
        T, Tdo, TT, it = O.funk_time(T, Tdo, TT, it)
    sd, ud, bd = to_dev(s), to_dev(u), to_dev(b)
    s2, u2, b2 = torch.empty_like(sd), torch.empty_like(ud), torch.empty_like(bd)
    Td = torch.tensor([1e7], dtype=torch.float64, device=dev())
    Tdod = torch.tensor([1e-8], dtype=torch.float64, device=dev())
    TTd = torch.zeros(1, dtype=torch.float64, device=dev())
    itd = torch.zeros(1, dtype=torch.int32, device=dev())
    for k in range(20):
        fn = api.add2 if k < 10 else api.kernel_tvd
        fn(g, sd, ud, bd, s2, u2, b2, Td, Tdod, 0, method)
        sd, s2, ud, u2, bd, b2 = s2, sd, u2, ud, b2, bd
        api.funk_time(Td, Tdod, TTd, itd)
    gs, gu, gb = to_host(sd), to_host(ud), to_host(bd)
    assert abs(TTd.item() - TT) <= 1e-12 * TT and itd.item() == it
    assert np.max(np.abs(gs - hs) / np.abs(hs)) <= 1e-10
    assert np.max(np.abs(gu - hu)) <= 1e-10
    assert np.max(np.abs(gb - hb)) <= 1e-10
