"""GPU: the one-pass SSP-RK2 step (gd2_ssprk2_step) against two stage launches.

The one-pass kernel keeps the intermediate state on chip; its contract is
bitwise equality with stage(U -> S) followed by stage(S, base U -> U', a = b =
1/2), including the wave-speed fold and the clamp counts. The stage kernels
are checked against the oracle in test_gpu_euler.py; here the shapes are
chosen so that strips of every height (the launch uses short strips on small
grids), the first and last strips, columns split across waves and grids
narrower than one wave all occur, on every boundary kind.
"""
import numpy as np
import pytest
import torch

from golikov_2d_amd import api
from golikov_2d_amd import problems as PB
from oracle import oracle as O
from tests.gpu_util import dev, grids, rel_l1, say, to_dev, to_host

pytestmark = pytest.mark.gpu

BCS = {
    "transmissive": (O.BC_TRANSMISSIVE,) * 4,
    "periodic": (O.BC_PERIODIC,) * 4,
    "reflective": (O.BC_REFLECTIVE,) * 4,
    "refl_x_per_y": (O.BC_REFLECTIVE, O.BC_TRANSMISSIVE, O.BC_PERIODIC, O.BC_PERIODIC),
    "per_x_mixed_y": (O.BC_PERIODIC, O.BC_PERIODIC, O.BC_TRANSMISSIVE, O.BC_REFLECTIVE),
    "mixed_y": (O.BC_TRANSMISSIVE, O.BC_REFLECTIVE, O.BC_REFLECTIVE, O.BC_TRANSMISSIVE),
}


def two_stage(g, sc, U, dt, want_stage=False):
    S = torch.empty_like(U)
    out = torch.empty_like(U)
    w = torch.zeros(1, dtype=torch.int64, device=dev())
    api.stage(g, sc, U, None, S, 0.0, 1.0, dt)
    api.stage(g, sc, S, U, out, 0.5, 0.5, dt, wmax_bits=w)
    return (out, w, S) if want_stage else (out, w)


def one_pass(g, sc, U, dt, want_stage=False):
    out = torch.empty_like(U)
    w = torch.zeros(1, dtype=torch.int64, device=dev())
    S = torch.full_like(U, float("nan")) if want_stage else None
    api.ssprk2_step(g, sc, U, out, dt, wmax_bits=w, ustage=S)
    return (out, w, S) if want_stage else (out, w)


@pytest.mark.parametrize("shape", [(130, 97), (63, 64), (64, 63), (2, 5), (200, 3), (59, 2), (118, 41),
                                   (60, 9), (177, 13)])
@pytest.mark.parametrize("bc", list(BCS))
@pytest.mark.parametrize("flux,order", [(api.FLUX_HLLC, 2), (api.FLUX_HLLC, 1), (api.FLUX_EXACT, 2),
                                        (api.FLUX_EXACT, 1), (api.FLUX_HLL, 2), (api.FLUX_HLLC_KOROLKOV, 2),
                                        (api.FLUX_HLLC_2D, 2), (api.FLUX_HLLC_2D, 1)])
def test_one_pass_step_equals_two_stages(shape, bc, flux, order):
    nx, ny = shape
    say(f"one-pass step {nx}x{ny} bc={bc} flux={flux} order={order}")
    g, og = grids(nx, ny, BCS[bc])
    sc = api.make_scheme(api.EULER, flux, order, api.SSPRK2, 0.4)
    U = to_dev(PB.four_quadrant(og, seed=nx * 7 + ny))
    dt = 0.3 / (api.wavespeed_max(g, sc, U))
    ref, wref, Sref = two_stage(g, sc, U, dt, want_stage=True)
    got, w = one_pass(g, sc, U, dt)
    torch.cuda.synchronize()
    assert torch.equal(got, ref), float((got - ref).abs().max())
    assert int(w.item()) == int(wref.item())
    # the variant that also stores the stage-1 state: every cell written once,
    # bitwise the first stage launch's output
    got2, w2, S = one_pass(g, sc, U, dt, want_stage=True)
    assert torch.equal(got2, ref) and int(w2.item()) == int(wref.item())
    assert torch.equal(S, Sref), float((S - Sref).abs().nan_to_num(1e300).max())


def test_one_pass_step_matches_oracle():
    say("one-pass step vs oracle 150x70")
    g, og = grids(150, 70, BCS["refl_x_per_y"])
    osc = O.Scheme(O.PHYS_EULER, O.FLUX_HLLC, 2, 0.4)
    U0 = PB.four_quadrant(og, seed=21)
    dt = 1e-3
    U1 = O.stage(og, osc, U0, None, 0.0, 1.0, dt)
    ref = O.stage(og, osc, U1, U0, 0.5, 0.5, dt)
    got, _ = one_pass(g, api.make_scheme(), to_dev(U0), dt)
    assert np.all(rel_l1(to_host(got), ref) <= 1e-12)


def test_one_pass_clamp_counts_equal_two_stages():
    """A near-vacuum expansion stepped with too large a dt fires the clamps;
    both paths count every clamped cell once."""
    say("clamp counts")
    g, og = grids(150, 90, BCS["transmissive"])
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, 2, api.SSPRK2, 0.4)
    U = PB.uniform_euler(og, rho=1.0, p=1e-4, u=0.0, v=0.0)
    U[1] = np.where(np.arange(150)[None, :] < 75, -2.0, 2.0) * U[0]   # rho u: pulled apart
    U[3] = 1e-4 / 0.4 + 0.5 * U[1] ** 2 / U[0]
    Ud = to_dev(U)
    dt = 3.0 / api.wavespeed_max(g, sc, Ud)
    api.diagnostics(reset=True)
    ref, wref = two_stage(g, sc, Ud, dt)
    d_ref = api.diagnostics(reset=True)
    got, w = one_pass(g, sc, Ud, dt)
    d = api.diagnostics(reset=True)
    assert d_ref["clamp_rho"] + d_ref["clamp_p"] > 0
    assert (d["clamp_rho"], d["clamp_p"]) == (d_ref["clamp_rho"], d_ref["clamp_p"])
    assert torch.equal(got, ref) and int(w.item()) == int(wref.item())
    got2, _, S = one_pass(g, sc, Ud, dt, want_stage=True)
    d2 = api.diagnostics(reset=True)
    assert (d2["clamp_rho"], d2["clamp_p"]) == (d_ref["clamp_rho"], d_ref["clamp_p"])
    assert torch.equal(got2, ref)


@pytest.mark.parametrize("nsteps", [1, 4, 7])
def test_solver_context_one_pass_equals_two_pass(nsteps):
    """gd2_ctx_advance ping-pongs one-pass steps between the state and its
    scratch (copying back after an odd count): same state, clock and steps
    as stage-by-stage stepping."""
    say(f"ctx one-pass vs two-pass, {nsteps} steps")
    g, og = grids(257, 130, BCS["transmissive"])
    sc = api.make_scheme()
    U0 = to_dev(PB.four_quadrant(og, seed=4))
    # reference: explicit two-stage steps with the context's CFL rule (the
    # first bound from the state, later ones from the previous step's fold)
    Ur = U0.clone()
    t = 0.0
    wb = api.wavespeed_max(g, sc, Ur)
    for _ in range(nsteps):
        dt = sc.cfl / wb
        Ur, w = two_stage(g, sc, Ur, dt)
        wb = api.bits_to_double(w)
        t += dt
    prev = api.set_step_path(api.STEP_ONE_PASS)  # the grid is below the size where AUTO picks it
    try:
        assert api.one_pass_preferred(g, sc)
        s = api.Solver(g, sc)
        U = U0.clone()
        s.start(U)
        s.advance(U, nsteps)
        tg, ng, _ = s.time()
        s.close()
    finally:
        api.set_step_path(prev)
    assert ng == nsteps
    assert torch.equal(U, Ur)
    assert tg == t


def test_fused_clock_steps_mix_with_clock_kernel_steps():
    """The one-pass step with no end time updates the device clock inside the
    step kernel; steps with an end time go through the clock kernel. Mixed
    in any order (each switch moves between the two three-slot schemes),
    they give the states, clocks and dt logs of stage-by-stage stepping, bit
    for bit, including the no-op steps after the end time."""
    say("fused clock vs clock kernel")
    g, og = grids(300, 170, BCS["refl_x_per_y"])
    sc = api.make_scheme()
    U0 = to_dev(PB.four_quadrant(og, seed=11))
    res = []
    for mode in (api.STEP_ONE_PASS, api.STEP_TWO_PASS):
        prev = api.set_step_path(mode)
        try:
            s = api.Solver(g, sc)
            log = torch.full((40,), -1.0, dtype=torch.float64, device=dev())
            s.log_dt(log)
            U = U0.clone()
            s.start(U)
            s.advance(U, 3)                  # fused (one-pass)
            t3, _, _ = s.time()
            s.advance(U, 2, t_end=1.0)       # clock kernel
            s.advance(U, 4)                  # fused
            t9, n9, _ = s.time()
            s.advance(U, 3, t_end=t9 + 0.5 * (t9 - t3) / 6)  # reaches t_end, then no-ops
            s.advance(U, 1)                  # fused
            res.append((U, s.time(), to_host(log)))
            s.close()
        finally:
            api.set_step_path(prev)
    (Ua, ca, la), (Ub, cb, lb) = res
    assert torch.equal(Ua, Ub)
    assert ca == cb and ca[1] == 9 + 1 + 1
    assert np.array_equal(la, lb) and np.all(la[:11] > 0) and np.all(la[11:] == -1.0)


@pytest.mark.parametrize("path", ["one_pass", "two_pass"])
def test_solver_kernel_timing_changes_nothing(path):
    """gd2_ctx_time_kernels / gd2_ctx_kernel_ms: events around every
    stride-th step's kernels up to `capacity` steps; the state and the clock
    are bitwise those of an untimed run; the count restarts after each read."""
    say(f"ctx kernel timing, {path}")
    g, og = grids(300, 200, BCS["transmissive"])
    sc = api.make_scheme()
    U0 = to_dev(PB.four_quadrant(og, seed=6))
    prev = api.set_step_path(api.STEP_ONE_PASS if path == "one_pass" else api.STEP_TWO_PASS)
    try:
        runs = []
        for timed in (False, True):
            s = api.Solver(g, sc)
            U = U0.clone()
            s.start(U)
            if timed:
                s.time_kernels(3, stride=3)
            s.advance(U, 8)
            if timed:
                ms, n = s.kernel_ms()
                assert n == 3 and ms > 0.0  # steps 0, 3, 6 of 8
                assert s.kernel_ms() == (0.0, 0)
                s.time_kernels(0)
                s.advance(U, 2)
                assert s.kernel_ms() == (0.0, 0)
            else:
                s.advance(U, 2)
            runs.append((U, s.time()))
            s.close()
    finally:
        api.set_step_path(prev)
    assert torch.equal(runs[0][0], runs[1][0])
    assert runs[0][1] == runs[1][1]


@pytest.mark.parametrize("path", ["one_pass", "two_pass"])
def test_16384_square_rows_equal_one_row_solution(path):
    """64-bit indexing at 16384^2 (8.6 GB per state, plane offsets beyond
    2^32 bytes): an x-directed state, constant along y, stays so under
    SSP-RK2 with transmissive sides, and every row of the 16384^2 run equals
    bit for bit the row of the same problem on 16384 x 4 (whose rows are
    checked against the oracle)."""
    n = 16384
    say(f"16384^2 rows, {path}")
    x = (np.arange(n) + 0.5) / n
    rho = 1.0 + 0.3 * np.sin(6 * np.pi * x) + np.where(x < 0.5, 0.5, 0.0)
    p = 1.0 + 0.2 * np.cos(4 * np.pi * x) + np.where(x < 0.3, 0.4, 0.0)
    u = 0.3 * np.sin(2 * np.pi * x)
    v = np.full(n, 0.2)
    row = PB.euler_from_prim(1.4, rho, p, u, v)           # (4, n)
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, 2, api.SSPRK2, 0.4)
    prev = api.set_step_path(api.STEP_ONE_PASS if path == "one_pass" else api.STEP_TWO_PASS)
    try:
        out = {}
        for ny in (4, n):
            g = api.make_grid(n, ny, dx=1.0 / n, dy=1.0 / n)
            U = to_dev(row[:, None, :]).repeat(1, ny, 1).contiguous()
            s = api.Solver(g, sc)
            s.start(U)
            s.advance(U, 3)
            out[ny] = (U, s.time())
            s.close()
        small, (t4, n4, _) = out[4]
        big, (tb, nb, _) = out[n]
        assert n4 == nb == 3 and t4 == tb
        for f in range(4):
            assert torch.equal(big[f], small[f, :1, :].expand(n, n)), f
        # the one-row problem against the oracle
        og = O.make_grid(n, 4, dx=1.0 / n, dy=1.0 / n)
        Uo = np.ascontiguousarray(np.repeat(row[:, None, :], 4, axis=1))
        steps, _ = O.run(og, O.Scheme(O.PHYS_EULER, O.FLUX_HLLC, 2, 0.4), 2, Uo, nsteps=3)
        assert steps == 3
        assert np.all(rel_l1(to_host(small), Uo) <= 1e-12)
    finally:
        api.set_step_path(prev)
