"""GPU parity of the Euler path through the C ABI against the CPU oracle.

Bar (per conserved field, relative L1 = sum|gpu - oracle| / sum|oracle|):
one stage or step from identical inputs <= 1e-12; 100 SSP-RK2 steps
<= 1e-10; density and pressure positive; no clamp hits. The face time step
of the reference-layout entry points and the CFL wave-speed maximum agree
to 1e-14 relative.
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

BCS = {
    "transmissive": (O.BC_TRANSMISSIVE,) * 4,
    "periodic": (O.BC_PERIODIC,) * 4,
    "mixed": (O.BC_REFLECTIVE, O.BC_TRANSMISSIVE, O.BC_PERIODIC, O.BC_PERIODIC),
}


def oscheme(flux, order, cfl=0.4):
    return O.Scheme(O.PHYS_EULER, flux, order, cfl)


def pscheme(flux, order, integ=api.SSPRK2, cfl=0.4):
    return api.make_scheme(api.EULER, flux, order, integ, cfl)


@pytest.mark.parametrize("shape", [(130, 97), (63, 64), (64, 63), (1, 5), (200, 3)])
@pytest.mark.parametrize("bc", list(BCS))
@pytest.mark.parametrize("flux,order", [(api.FLUX_HLLC, 2), (api.FLUX_HLLC, 1), (api.FLUX_EXACT, 1),
                                        (api.FLUX_EXACT, 2), (api.FLUX_HLL, 2), (api.FLUX_HLL, 1),
                                        (api.FLUX_HLLC_KOROLKOV, 2), (api.FLUX_HLLC_KOROLKOV, 1),
                                        (api.FLUX_HLLC_2D, 2), (api.FLUX_HLLC_2D, 1)])
def test_stage_matches_oracle(shape, bc, flux, order):
    nx, ny = shape
    say(f"stage {nx}x{ny} bc={bc} flux={flux} order={order}")
    g, og = grids(nx, ny, BCS[bc])
    U = PB.four_quadrant(og, seed=nx * 1000 + ny)
    dt = 0.2 / (nx + ny)
    ref, wref = O.stage(og, oscheme(flux, order), U, None, 0.0, 1.0, dt, want_w=True)
    Ud = to_dev(U)
    out = torch.empty_like(Ud)
    w = torch.zeros(1, dtype=torch.int64, device=dev())
    api.stage(g, pscheme(flux, order), Ud, None, out, 0.0, 1.0, dt, wmax_bits=w)
    got = to_host(out)
    assert np.all(rel_l1(got, ref) <= STEP_TOL), rel_l1(got, ref)
    assert abs(api.bits_to_double(w) - wref) <= 1e-14 * wref


@pytest.mark.parametrize("bc", list(BCS))
def test_ssprk2_second_stage_matches_oracle(bc):
    say(f"ssprk2 stage 2 bc={bc}")
    g, og = grids(150, 70, BCS[bc])
    U0 = PB.four_quadrant(og, seed=21)
    sc = oscheme(O.FLUX_HLLC, 2)
    dt = 1e-3
    U1 = O.stage(og, sc, U0, None, 0.0, 1.0, dt)
    ref = O.stage(og, sc, U1, U0, 0.5, 0.5, dt)
    U0d, U1d = to_dev(U0), to_dev(U1)
    dtd = torch.tensor([dt], dtype=torch.float64, device=dev())
    api.stage(g, pscheme(api.FLUX_HLLC, 2), U1d, U0d, U0d, 0.5, 0.5, dt_dev=dtd)  # in place on the base
    assert np.all(rel_l1(to_host(U0d), ref) <= STEP_TOL)


@pytest.mark.parametrize("n", [128, 160])
def test_100_step_ssprk2_run_matches_oracle(n):
    say(f"100-step run {n}^2")
    g, og = grids(n, n, BCS["transmissive"])
    U = PB.four_quadrant(og, seed=5)
    Ud = to_dev(U)
    O.diag_reset()
    steps, t = O.run(og, oscheme(O.FLUX_HLLC, 2), 2, U, nsteps=100)
    api.diagnostics(reset=True)
    s = api.Solver(g, pscheme(api.FLUX_HLLC, 2))
    s.start(Ud)
    s.advance(Ud, 100)
    tg, ng, _ = s.time()
    got = to_host(Ud)
    assert ng == steps == 100
    assert abs(tg - t) <= 1e-12 * t
    assert np.all(rel_l1(got, U) <= RUN_TOL), rel_l1(got, U)
    rho = got[0]
    p = 0.4 * (got[3] - 0.5 * (got[1] ** 2 + got[2] ** 2) / rho)
    assert rho.min() > 0 and p.min() > 0
    d = api.diagnostics()
    assert d["clamp_rho"] == 0 and d["clamp_p"] == 0
    s.close()


def test_sod_godunov_to_t02_matches_oracle():
    say("C0 Sod 200x200 exact Godunov to t=0.2")
    g, og = grids(200, 200, BCS["transmissive"])
    U = PB.sod_x(og)
    Ud = to_dev(U)
    steps, t = O.run(og, oscheme(O.FLUX_EXACT, 1), 1, U, t_end=0.2)
    s = api.Solver(g, pscheme(api.FLUX_EXACT, 1, api.FORWARD_EULER))
    s.start(Ud)
    tg, ng = s.run_to(Ud, 0.2, chunk=64)
    got = to_host(Ud)
    assert ng == steps and tg == t == 0.2
    assert np.all(rel_l1(got, U) <= RUN_TOL), rel_l1(got, U)
    assert np.all(got == got[:, :1, :])  # y-invariant
    d = api.diagnostics()
    assert d["newton_cap_hits"] == 0 and d["bisect_cap_hits"] == 0
    s.close()


@pytest.mark.parametrize("kind,order", [("ker_dekard", 1), ("kernel_tvd_old", 2)])
@pytest.mark.parametrize("method", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("bc", [(O.BC_TRANSMISSIVE, O.BC_FROZEN, O.BC_TRANSMISSIVE, O.BC_FROZEN),
                                (O.BC_REFLECTIVE, O.BC_PERIODIC, O.BC_TRANSMISSIVE, O.BC_TRANSMISSIVE)])
def test_reference_layout_step_matches_oracle(kind, order, method, bc):
    if bc[0] == O.BC_REFLECTIVE and bc[1] == O.BC_PERIODIC:
        bc = (O.BC_REFLECTIVE, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE, O.BC_TRANSMISSIVE)
    say(f"reference layout {kind} method={method} bc={bc}")
    g, og = grids(131, 77, bc, gamma=5.0 / 3.0, krit=0.25)
    U = PB.four_quadrant(og, seed=8)
    s, u, _ = O.cons_to_prim(og, O.PHYS_EULER, U)
    T0, Tdo = 1e7, 3e-4
    s2r, u2r, Tr = O.ref_euler_step(kind, og, s, u, T0, Tdo, method)
    sd, ud = to_dev(s), to_dev(u)
    s2, u2 = torch.empty_like(sd), torch.empty_like(ud)
    T = torch.tensor([T0], dtype=torch.float64, device=dev())
    T_do = torch.tensor([Tdo], dtype=torch.float64, device=dev())
    fn = api.ker_dekard if kind == "ker_dekard" else api.kernel_tvd_old
    fn(g, sd, ud, s2, u2, T, T_do, method)
    s2g, u2g = to_host(s2), to_host(u2)
    assert np.max(np.abs(s2g - s2r) / np.abs(s2r)) <= 1e-12
    assert np.max(np.abs(u2g - u2r)) <= 1e-12
    assert abs(T.item() - Tr) <= 1e-14 * Tr


def test_funk_time_rotation():
    say("funk_time")
    T = torch.tensor([0.01], dtype=torch.float64, device=dev())
    T_do = torch.tensor([1e-8], dtype=torch.float64, device=dev())
    TT = torch.tensor([0.5], dtype=torch.float64, device=dev())
    i = torch.tensor([4], dtype=torch.int32, device=dev())
    api.funk_time(T, T_do, TT, i)
    torch.cuda.synchronize()
    assert (T.item(), T_do.item(), TT.item(), i.item()) == O.funk_time(0.01, 1e-8, 0.5, 4)


def test_layout_conversion_roundtrip():
    say("prim<->cons")
    g, og = grids(70, 40, BCS["periodic"])
    U = PB.four_quadrant(og, seed=3)
    s, u, _ = O.cons_to_prim(og, O.PHYS_EULER, U)
    sd, ud = to_dev(s), to_dev(u)
    Ud = api.prim_to_cons(g, api.EULER, sd, ud)
    assert np.max(np.abs(to_host(Ud) - O.prim_to_cons(og, O.PHYS_EULER, s, u))) <= 1e-15 * 4
    s2, u2, _ = api.cons_to_prim(g, api.EULER, to_dev(U))
    assert np.array_equal(to_host(s2), s) and np.array_equal(to_host(u2), u)


def test_full_size_stage_parity_4096():
    """One HLLC-MUSCL stage at the 4096^2 benchmark size against the oracle."""
    say("4096^2 stage parity")
    n = 4096
    g, og = grids(n, n, BCS["transmissive"])
    U = PB.four_quadrant(og, seed=0)
    dt = 0.4 / O.wavespeed_max(og, oscheme(O.FLUX_HLLC, 2), U)
    ref = O.stage(og, oscheme(O.FLUX_HLLC, 2), U, None, 0.0, 1.0, dt)
    Ud = to_dev(U)
    out = torch.empty_like(Ud)
    api.stage(g, pscheme(api.FLUX_HLLC, 2), Ud, None, out, 0.0, 1.0, dt)
    got = to_host(out)
    assert np.all(rel_l1(got, ref) <= STEP_TOL), rel_l1(got, ref)


def test_full_size_100_steps_4096_match_oracle():
    """The benchmarked workload itself: 100 SSP-RK2 steps of HLLC-MUSCL at
    4096^2 (seed 0) through the solver context (the one-pass kernel, with its
    own clock update) against 100 oracle steps from the same state: relative
    L1 per field <= 1e-10, rho and p positive, no clamps, same step count and
    time."""
    n = 4096
    say("4096^2 100 steps vs oracle")
    g, og = grids(n, n, BCS["transmissive"])
    sc = pscheme(api.FLUX_HLLC, 2)
    assert api.one_pass_preferred(g, sc)
    U = PB.four_quadrant(og, seed=0)
    Ud = to_dev(U)
    api.diagnostics(reset=True)
    s = api.Solver(g, sc)
    s.start(Ud)
    s.advance(Ud, 100)
    t_gpu, n_gpu, _ = s.time()
    s.close()
    d = api.diagnostics(reset=True)
    got = to_host(Ud)
    del Ud
    steps, t_ref = O.run(og, oscheme(O.FLUX_HLLC, 2), 2, U, nsteps=100)
    err = rel_l1(got, U)
    say(f"  rel L1 per field {err}")
    assert steps == n_gpu == 100
    assert abs(t_gpu - t_ref) <= 1e-12 * t_ref
    assert np.all(err <= 1e-10), err
    rho = got[0]
    p = 0.4 * (got[3] - 0.5 * (got[1] ** 2 + got[2] ** 2) / rho)
    assert rho.min() > 0 and p.min() > 0
    assert d["clamp_rho"] == 0 and d["clamp_p"] == 0


def test_full_size_invariants_4096():
    """Size-independent properties at 4096^2: free stream is preserved exactly,
    periodic totals are conserved, x-directed data stays y-invariant."""
    n = 4096
    say("4096^2 invariants")
    g = api.make_grid(n, n, bc=(api.BC_PERIODIC,) * 4)
    og = O.make_grid(n, n, bc=(O.BC_PERIODIC,) * 4)
    U = to_dev(PB.uniform_euler(og, rho=0.8, p=1.1, u=0.3, v=-0.6))
    s = api.Solver(g, pscheme(api.FLUX_HLLC, 2))
    s.start(U)
    U0 = U.clone()
    s.advance(U, 3)
    assert torch.equal(U, U0)
    U = to_dev(PB.four_quadrant(og, seed=1))
    tot0 = U.sum(dim=(1, 2))
    s.start(U)
    s.advance(U, 5)
    tot1 = U.sum(dim=(1, 2))
    assert torch.all((tot1 - tot0).abs() <= 1e-12 * U.abs().sum(dim=(1, 2)))
    s.close()
    g2 = api.make_grid(n, 64)
    og2 = O.make_grid(n, 64)
    U = to_dev(PB.sod_x(og2))
    s2 = api.Solver(g2, pscheme(api.FLUX_HLLC, 2))
    s2.start(U)
    s2.advance(U, 10)
    assert torch.equal(U, U[:, :1, :].expand_as(U))
    s2.close()


@pytest.mark.parametrize("method", [0, 1, 2, 3, 4])
def test_kernel_tvd_old_axisymmetric_sphere_matches_oracle(method):
    """Kernel_TVD_old's 1/y source terms and inner sphere (kernel.cu:3303,
    3651-3662) with its frozen east side and axis (reflective) south side:
    one step against the oracle, 1e-12."""
    say(f"kernel_tvd_old axisymmetric method={method}")
    bc = (O.BC_TRANSMISSIVE, O.BC_FROZEN, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE)
    g, og = grids(157, 90, bc, gamma=1.4, krit=0.25, r_frozen=0.23, axisym=True)
    U = PB.four_quadrant(og, seed=11)
    s, u, _ = O.cons_to_prim(og, O.PHYS_EULER, U)
    T0, Tdo = 1e7, 2e-4
    s2r, u2r, Tr = O.ref_euler_step("kernel_tvd_old", og, s, u, T0, Tdo, method)
    sd, ud = to_dev(s), to_dev(u)
    s2, u2 = torch.empty_like(sd), torch.empty_like(ud)
    T = torch.tensor([T0], dtype=torch.float64, device=dev())
    T_do = torch.tensor([Tdo], dtype=torch.float64, device=dev())
    api.kernel_tvd_old(g, sd, ud, s2, u2, T, T_do, method)
    s2g, u2g = to_host(s2), to_host(u2)
    assert np.max(np.abs(s2g - s2r) / np.abs(s2r)) <= 1e-12
    assert np.max(np.abs(u2g - u2r)) <= 1e-12
    assert abs(T.item() - Tr) <= 1e-14 * Tr
    x = og.x0 + np.arange(og.nx) * og.dx
    y = og.y0 + np.arange(og.ny) * og.dy
    X, Y = np.meshgrid(x, y)
    inside = np.sqrt(X * X + Y * Y) < og.r_frozen
    assert inside.sum() > 20
    np.testing.assert_array_equal(s2g[inside], s[inside])
    # Ker_Dekard takes neither option
    with pytest.raises(api.GD2Error):
        api.ker_dekard(g, sd, ud, s2, u2, T, T_do, method)


@pytest.mark.parametrize("shape", [(1, 5), (5, 1), (2, 2), (3, 4), (66, 3)])
def test_kernel_tvd_old_options_small_grids_match_oracle(shape):
    """Kernel_TVD_old's sphere and 1/y terms on grids a few cells wide."""
    nx, ny = shape
    say(f"kernel_tvd_old options small grid {nx}x{ny}")
    bc = (O.BC_TRANSMISSIVE, O.BC_FROZEN, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE)
    g, og = grids(nx, ny, bc, gamma=1.4, dx=0.05, dy=0.05, x0=0.025, y0=0.025, r_frozen=0.06, axisym=True)
    rng = np.random.default_rng(nx * 13 + ny)
    s = np.ascontiguousarray(np.stack([rng.uniform(0.5, 1.5, (ny, nx)), rng.uniform(0.5, 1.5, (ny, nx))], -1))
    u = np.ascontiguousarray(rng.uniform(-0.5, 0.5, (ny, nx, 2)))
    for method in (0, 1, 2):
        s2r, u2r, Tr = O.ref_euler_step("kernel_tvd_old", og, s, u, 1e7, 1e-4, method)
        sd, ud = to_dev(s), to_dev(u)
        s2, u2 = torch.empty_like(sd), torch.empty_like(ud)
        T = torch.tensor([1e7], dtype=torch.float64, device=dev())
        T_do = torch.tensor([1e-4], dtype=torch.float64, device=dev())
        api.kernel_tvd_old(g, sd, ud, s2, u2, T, T_do, method)
        assert np.max(np.abs(to_host(s2) - s2r) / np.abs(s2r)) <= 1e-12, method
        assert np.max(np.abs(to_host(u2) - u2r)) <= 1e-12, method
        assert T.item() == Tr or abs(T.item() - Tr) <= 1e-14 * Tr, method


@pytest.mark.parametrize("method", [0, 1, 2])
def test_kernel_tvd_old_rows_below_the_axis_match_oracle(method):
    """Kernel_TVD_old's 1/y terms take the signed y (kernel.cu:3291,
    3651-3662): a grid across the axis, one step against the oracle, 1e-12."""
    say(f"kernel_tvd_old rows below the axis method={method}")
    bc = (O.BC_TRANSMISSIVE, O.BC_FROZEN, O.BC_TRANSMISSIVE, O.BC_TRANSMISSIVE)
    h = 0.01
    g, og = grids(140, 77, bc, gamma=1.4, krit=0.25, dx=h, dy=h, x0=h / 2, y0=-40.5 * h, axisym=True)
    U = PB.four_quadrant(og, seed=13)
    s, u, _ = O.cons_to_prim(og, O.PHYS_EULER, U)
    T0, Tdo = 1e7, 2e-4
    s2r, u2r, Tr = O.ref_euler_step("kernel_tvd_old", og, s, u, T0, Tdo, method)
    sd, ud = to_dev(s), to_dev(u)
    s2, u2 = torch.empty_like(sd), torch.empty_like(ud)
    T = torch.tensor([T0], dtype=torch.float64, device=dev())
    T_do = torch.tensor([Tdo], dtype=torch.float64, device=dev())
    api.kernel_tvd_old(g, sd, ud, s2, u2, T, T_do, method)
    assert np.max(np.abs(to_host(s2) - s2r) / np.abs(s2r)) <= 1e-12
    assert np.max(np.abs(to_host(u2) - u2r)) <= 1e-12
    assert abs(T.item() - Tr) <= 1e-14 * Tr


def _west_inflow_rows(ny, nx, seed):
    """Random state whose west edge cells are subsonic with vx > -5 (the
    ghost's rule acts), supersonic, subsonic with vx <= -5, or supersonic at
    their own pressure and subsonic at 1/(gamma M_inf^2) (M_inf 0.7, gamma 1.4)."""
    rng = np.random.default_rng(seed)
    s = np.stack([rng.uniform(0.6, 1.4, (ny, nx)), rng.uniform(0.6, 1.4, (ny, nx))], -1)
    u = rng.uniform(-0.4, 0.4, (ny, nx, 2))
    u[1::7, 0, 0] = 3.0
    u[2::7, :2, 0] = -6.0
    s[2::7, :2, 1] = 80.0
    s[3::7, 0] = (1.0, 0.6)
    u[3::7, 0] = (1.3, 0.0)
    return np.ascontiguousarray(s), np.ascontiguousarray(u)


def _tvd_old_dev(g, s, u, T0, Tdo, method):
    sd, ud = to_dev(s), to_dev(u)
    s2, u2 = torch.empty_like(sd), torch.empty_like(ud)
    T = torch.tensor([T0], dtype=torch.float64, device=dev())
    T_do = torch.tensor([Tdo], dtype=torch.float64, device=dev())
    api.kernel_tvd_old(g, sd, ud, s2, u2, T, T_do, method)
    return to_host(s2), to_host(u2), T.item()


@pytest.mark.parametrize("method", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("options", [False, True])
def test_kernel_tvd_old_west_inflow_matches_oracle(method, options):
    """Kernel_TVD_old with the reference's own sides: the west inflow ghost
    (kernel.cu:3323-3337, 3365-3378), frozen east column, axis (reflective)
    south, transmissive north; with and without its sphere and 1/y terms.
    One step against the oracle, 1e-12, over three wave strips."""
    say(f"kernel_tvd_old west inflow method={method} options={options}")
    bc = (O.BC_WEST_INFLOW, O.BC_FROZEN, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE)
    kw = dict(r_frozen=0.23, axisym=True) if options else {}
    g, og = grids(157, 90, bc, gamma=1.4, krit=0.25, m_inf=0.7, **kw)
    s, u = _west_inflow_rows(90, 157, seed=method)
    T0, Tdo = 1e7, 2e-4
    s2r, u2r, Tr = O.ref_euler_step("kernel_tvd_old", og, s, u, T0, Tdo, method)
    s2, u2, T = _tvd_old_dev(g, s, u, T0, Tdo, method)
    assert np.max(np.abs(s2 - s2r) / np.abs(s2r)) <= 1e-12
    assert np.max(np.abs(u2 - u2r)) <= 1e-12
    assert abs(T - Tr) <= 1e-14 * Tr
    # the west side acts: the step differs from the transmissive one near it
    gt, ogt = grids(157, 90, (O.BC_TRANSMISSIVE,) + bc[1:], gamma=1.4, krit=0.25, **kw)
    s2t, _, _ = _tvd_old_dev(gt, s, u, T0, Tdo, method)
    assert np.any(s2t[:, 0] != s2[:, 0]) and np.any(s2t[:, 1] != s2[:, 1])
    np.testing.assert_array_equal(s2t[:, 2:], s2[:, 2:])


@pytest.mark.parametrize("shape", [(1, 5), (2, 3), (3, 4), (64, 9), (66, 3)])
def test_kernel_tvd_old_west_inflow_small_grids_match_oracle(shape):
    """The west inflow ghost on grids one to a few cells wide (column 1 missing
    or the east ghost), transmissive east."""
    nx, ny = shape
    say(f"kernel_tvd_old west inflow small grid {nx}x{ny}")
    bc = (O.BC_WEST_INFLOW, O.BC_TRANSMISSIVE, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE)
    g, og = grids(nx, ny, bc, gamma=1.4, dx=0.05, dy=0.05, m_inf=0.7)
    s, u = _west_inflow_rows(ny, nx, seed=nx * 13 + ny)
    for method in (0, 1, 2):
        s2r, u2r, Tr = O.ref_euler_step("kernel_tvd_old", og, s, u, 1e7, 1e-4, method)
        s2, u2, T = _tvd_old_dev(g, s, u, 1e7, 1e-4, method)
        assert np.max(np.abs(s2 - s2r) / np.abs(s2r)) <= 1e-12, method
        assert np.max(np.abs(u2 - u2r)) <= 1e-12, method
        assert abs(T - Tr) <= 1e-14 * Tr, method


def test_west_inflow_rejected_elsewhere():
    """WEST_INFLOW: Kernel_TVD_old only, west side only, M_inf > 0."""
    say("west inflow rejected elsewhere")
    s, u = _west_inflow_rows(5, 6, seed=0)
    ok = (O.BC_WEST_INFLOW, O.BC_FROZEN, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE)
    _tvd_old_dev(api.make_grid(6, 5, bc=ok, m_inf=0.5), s, u, 1e7, 1e-4, 1)
    east = (O.BC_TRANSMISSIVE, O.BC_WEST_INFLOW, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE)
    for g in (api.make_grid(6, 5, bc=ok, m_inf=0.0), api.make_grid(6, 5, bc=east, m_inf=0.5)):
        with pytest.raises(api.GD2Error):
            _tvd_old_dev(g, s, u, 1e7, 1e-4, 1)
    sd, ud = to_dev(s), to_dev(u)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev())
    with pytest.raises(api.GD2Error):
        api.ker_dekard(api.make_grid(6, 5, bc=ok, m_inf=0.5), sd, ud, torch.empty_like(sd), torch.empty_like(ud),
                       T, T.clone(), 1)


@pytest.mark.parametrize("method", [0, 1, 2])
def test_kernel_tvd_old_reference_sides_30_steps_match_oracle(method):
    """Kernel_TVD_old driven as the reference's step loop drives its steps
    (lagged face time step, funk_time after each, ping-pong buffers) with its
    own sides: west inflow ghost (M_inf 0.7), frozen east column, axis
    south, transmissive north, inner sphere and 1/y terms; 30 steps of a
    smooth subsonic state against the oracle, 1e-10."""
    nx, ny = 190, 120
    say(f"kernel_tvd_old reference sides 30 steps method={method}")
    bc = (O.BC_WEST_INFLOW, O.BC_FROZEN, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE)
    g, og = grids(nx, ny, bc, gamma=1.4, krit=0.25, dx=1.0 / nx, dy=0.6 / ny, m_inf=0.7, r_frozen=0.12,
                  axisym=True)
    x = og.x0 + np.arange(nx) * og.dx
    y = og.y0 + np.arange(ny) * og.dy
    X, Y = np.meshgrid(x, y)
    s = np.empty((ny, nx, 2))
    u = np.empty((ny, nx, 2))
    s[..., 0] = 1.0 + 0.2 * np.sin(6 * X) * np.cos(5 * Y)
    s[..., 1] = 1.0 / (1.4 * 0.49) * (1 + 0.1 * np.cos(4 * X + 3 * Y))
    u[..., 0] = -0.3 + 0.2 * np.sin(3 * Y)
    u[..., 1] = 0.1 * np.sin(4 * X) * np.sin(3 * Y)
    s, u = np.ascontiguousarray(s), np.ascontiguousarray(u)
    hs, hu = s.copy(), u.copy()
    T, Tdo, TT, it = 1e7, 1e-8, 0.0, 0
    for _ in range(30):
        hs, hu, T = O.ref_euler_step("kernel_tvd_old", og, hs, hu, T, Tdo, method)
        T, Tdo, TT, it = O.funk_time(T, Tdo, TT, it)
    sd, ud = to_dev(s), to_dev(u)
    s2, u2 = torch.empty_like(sd), torch.empty_like(ud)
    Td = torch.tensor([1e7], dtype=torch.float64, device=dev())
    Tdod = torch.tensor([1e-8], dtype=torch.float64, device=dev())
    TTd = torch.zeros(1, dtype=torch.float64, device=dev())
    itd = torch.zeros(1, dtype=torch.int32, device=dev())
    for _ in range(30):
        api.kernel_tvd_old(g, sd, ud, s2, u2, Td, Tdod, method)
        sd, s2, ud, u2 = s2, sd, u2, ud
        api.funk_time(Td, Tdod, TTd, itd)
    gs, gu = to_host(sd), to_host(ud)
    assert abs(TTd.item() - TT) <= 1e-12 * TT and itd.item() == it == 30
    assert np.max(np.abs(gs - hs) / np.abs(hs)) <= 1e-10
    assert np.max(np.abs(gu - hu)) <= 1e-10
    assert np.all(gs[..., 0] > 0) and np.all(gs[..., 1] > 0)
