"""GPU parity of the Godunov-Kolgan-Rodionov step (gd2_kernel_gkr, the
reference's Kernel_GKR, kernel.cu:2663) with the oracle.

One step on the perturbed stellar-wind state of the production grid at two
sizes, both fluxes (HLLD, the reference's, and HLLC-MHD), with the
reference's frozen radius 2.0 and a small one; tiny grids where the side
rules and the second-neighbour choices of the first and last two rows and
columns overlap, and grids whose interior (handled by the march, the
two-cell band by the per-cell kernel) is one cell, one row or column, or
several strips wide and high; a 20-step loop with gd2_funk_time against the
oracle's loop; clamp counts equal. Bar: relative L1 per field 1e-12 for one step, 1e-10
after 20; the face time step to 1e-14.
"""
import numpy as np
import pytest
import torch

from golikov_2d_amd import api
from golikov_2d_amd import problems as PB
from oracle import oracle as O
from tests.gpu_util import dev, say, to_dev, to_host

pytestmark = pytest.mark.gpu


def wind_state(nx, ny, r_frozen, seed):
    kw = dict(PB.wind_grid_args(nx, ny), r_frozen=r_frozen)
    g, og = api.make_grid(nx, ny, **kw), O.make_grid(nx, ny, **kw)
    s, u, b = PB.stellar_wind_prim(og)
    rng = np.random.default_rng(seed)
    s = s * (1.0 + 0.05 * rng.random(s.shape))
    u = u + 0.05 * rng.standard_normal(u.shape)
    b = b + 0.05 * rng.standard_normal(b.shape)
    return g, og, np.ascontiguousarray(s), np.ascontiguousarray(u), np.ascontiguousarray(b)


def rel_l1_fields(got, ref):
    a = got.reshape(-1, got.shape[-1])
    b = ref.reshape(-1, ref.shape[-1])
    return np.abs(a - b).sum(0) / np.maximum(np.abs(b).sum(0), 1e-300)


def gpu_step(g, s, u, b, T0, Tdo, method):
    sd, ud, bd = to_dev(s), to_dev(u), to_dev(b)
    s2, u2, b2 = torch.empty_like(sd), torch.empty_like(ud), torch.empty_like(bd)
    T = torch.tensor([T0], dtype=torch.float64, device=dev())
    T_do = torch.tensor([Tdo], dtype=torch.float64, device=dev())
    api.kernel_gkr(g, sd, ud, bd, s2, u2, b2, T, T_do, 0, method)
    return to_host(s2), to_host(u2), to_host(b2), T.item()


@pytest.mark.parametrize("method", [3, 2])
@pytest.mark.parametrize("nx,ny,r_frozen", [(203, 131, 2.0), (322, 179, 0.3), (64, 40, 2.0)])
def test_gkr_step_matches_oracle(method, nx, ny, r_frozen):
    say(f"gkr step {nx}x{ny} r_frozen={r_frozen} method={method}")
    g, og, s, u, b = wind_state(nx, ny, r_frozen, nx + method)
    Tdo = 0.2 * min(og.dx, og.dy) / 4.0
    O.diag_reset()
    s2r, u2r, b2r, Tr = O.ref_gkr_step(og, s, u, b, 1e7, Tdo, method)
    dref = O.diag()
    api.diagnostics(reset=True)
    s2, u2, b2, T = gpu_step(g, s, u, b, 1e7, Tdo, method)
    d = api.diagnostics(reset=True)
    for got, ref in ((s2, s2r), (u2, u2r), (b2, b2r)):
        assert np.all(rel_l1_fields(got, ref) <= 1e-12), rel_l1_fields(got, ref)
    assert abs(T - Tr) <= 1e-14 * Tr
    assert (d["clamp_rho"], d["clamp_p"]) == (dref["clamp_rho"], dref["clamp_p"])


@pytest.mark.parametrize("shape", [(1, 1), (1, 4), (4, 1), (2, 2), (3, 5), (5, 3), (5, 5), (6, 6), (5, 9), (65, 5),
                                   (125, 70)])
def test_gkr_tiny_grids_match_oracle(shape):
    nx, ny = shape
    say(f"gkr tiny {nx}x{ny}")
    g, og, s, u, b = wind_state(nx, ny, 0.0, 7 * nx + ny)
    s2r, u2r, b2r, Tr = O.ref_gkr_step(og, s, u, b, 1e7, 1e-3, 3)
    s2, u2, b2, T = gpu_step(g, s, u, b, 1e7, 1e-3, 3)
    for got, ref in ((s2, s2r), (u2, u2r), (b2, b2r)):
        assert np.max(np.abs(got - ref) / (np.abs(ref) + 1e-3)) <= 1e-12
    assert abs(T - Tr) <= 1e-14 * Tr


def test_gkr_loop_with_funk_time_matches_oracle():
    """20 steps of gd2_kernel_gkr + gd2_funk_time (lagged face time step, the
    first step 1e-8) against the oracle's loop."""
    say("gkr 20-step loop")
    g, og, s, u, b = wind_state(161, 90, 2.0, 3)
    sr, ur, br = s.copy(), u.copy(), b.copy()
    T, Tdo, TT, i = 1e7, 1e-8, 0.0, 0
    for _ in range(20):
        sr, ur, br, T = O.ref_gkr_step(og, sr, ur, br, T, Tdo, 3)
        T, Tdo, TT, i = O.funk_time(T, Tdo, TT, i)
    sd, ud, bd = to_dev(s), to_dev(u), to_dev(b)
    s2, u2, b2 = torch.empty_like(sd), torch.empty_like(ud), torch.empty_like(bd)
    Td = torch.tensor([1e7], dtype=torch.float64, device=dev())
    Tdod = torch.tensor([1e-8], dtype=torch.float64, device=dev())
    TTd = torch.tensor([0.0], dtype=torch.float64, device=dev())
    idev = torch.tensor([0], dtype=torch.int32, device=dev())
    for k in range(20):
        if k % 2 == 0:
            api.kernel_gkr(g, sd, ud, bd, s2, u2, b2, Td, Tdod, 0, 3)
        else:
            api.kernel_gkr(g, s2, u2, b2, sd, ud, bd, Td, Tdod, 0, 3)
        api.funk_time(Td, Tdod, TTd, idev)
    for got, ref in ((to_host(sd), sr), (to_host(ud), ur), (to_host(bd), br)):
        assert np.all(rel_l1_fields(got, ref) <= 1e-10), rel_l1_fields(got, ref)
    assert abs(TTd.item() - TT) <= 1e-12 * TT and idev.item() == i == 20


def test_gkr_rejects_other_sides_and_fluxes():
    g, _, s, u, b = wind_state(16, 12, 0.0, 1)
    sd, ud, bd = to_dev(s), to_dev(u), to_dev(b)
    T = torch.tensor([1e7], dtype=torch.float64, device=dev())
    with pytest.raises(api.GD2Error):
        api.kernel_gkr(g, sd, ud, bd, sd.clone(), ud.clone(), bd.clone(), T, T.clone(), 0, 1)
    bad = api.make_grid(16, 12, **dict(PB.wind_grid_args(16, 12), bc=(api.BC_TRANSMISSIVE,) * 4))
    with pytest.raises(api.GD2Error):
        api.kernel_gkr(bad, sd, ud, bd, sd.clone(), ud.clone(), bd.clone(), T, T.clone(), 0, 3)
