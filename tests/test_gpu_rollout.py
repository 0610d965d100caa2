"""GPU checks of the differentiable rollout (LibTorch autograd Function over
the HIP stage and stage-VJP kernels, checkpoints every `every` steps).

* the rollout with a recorded dt schedule reproduces the native solver
  bitwise;
* gradients do not depend on the checkpoint interval (bitwise);
* the gradient of 1/2||rho(T) - rho_target||^2 over 200 steps at 128^2
  matches torch float64 autograd through the restated stages
  (tests/torch_ref.py) replaying the same dt schedule;
* MHD (HLLD / HLLC-MHD) rollouts of an Orszag-Tang state replay the native
  solver bitwise, and their gradients match autograd through the restated
  MHD stages and do not depend on the checkpoint interval;
* first-order rollouts at 1024^2 (200 steps) match central differences
  (h = 1e-6) at 64 random cells away from the initial discontinuities;
* second-order (minmod) rollouts at 1024^2 (200 steps): directional
  derivatives along smooth directions match central differences (h = 1e-6);
  single-cell differences are not used at this order: in smooth regions the
  two minmod arguments differ by O(dx^2) ~ h, so a one-cell perturbation
  crosses minmod kinks in its cone of influence and the quotient mixes
  branches.
"""
import numpy as np
import pytest
import torch

from golikov_2d_amd import adjoint as AD
from golikov_2d_amd import api
from golikov_2d_amd import problems as PB
from tests import torch_ref as TR
from tests.gpu_util import dev, say, to_dev

pytestmark = pytest.mark.gpu


def setup(n, order, seed=0, base_amp=1e-3, target_amp=1e-2, nsteps=200):
    g = api.make_grid(n, n)
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, order, api.SSPRK2, 0.4)
    quad = PB.four_quadrant(g, seed=seed)
    U0 = to_dev(PB.smooth_perturbed_euler(g, quad, seed + 100, base_amp) if base_amp > 0 else quad)
    Ut = to_dev(PB.smooth_perturbed_euler(g, quad, seed + 200, target_amp))
    dts, _ = AD.dt_schedule(g, sc, U0, nsteps)
    with torch.no_grad():
        rho_t = AD.rollout(g, sc, Ut, dts, every=nsteps + 1)[0].clone()
    return g, sc, U0, dts, rho_t


def grad_of_loss(g, sc, U0, dts, rho_t, every):
    X = U0.clone().requires_grad_(True)
    loss = AD.density_loss(AD.rollout(g, sc, X, dts, every=every), rho_t)
    loss.backward()
    torch.cuda.synchronize()
    return X.grad.detach().clone(), float(loss)


def test_rollout_replays_solver_bitwise():
    say("rollout vs native solver 96x80, 30 steps")
    g = api.make_grid(96, 80, bc=(api.BC_REFLECTIVE, api.BC_TRANSMISSIVE, api.BC_PERIODIC, api.BC_PERIODIC))
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, 2, api.SSPRK2, 0.4)
    U0 = to_dev(PB.four_quadrant(g, seed=3))
    dts, Usolver = AD.dt_schedule(g, sc, U0, 30)
    assert torch.all(dts > 0)
    with torch.no_grad():
        UT = AD.rollout(g, sc, U0, dts, every=7)
    assert torch.equal(UT, Usolver)


def test_gradient_independent_of_checkpoint_interval():
    say("checkpoint interval invariance 128^2, 45 steps")
    g, sc, U0, dts, rho_t = setup(128, 2, nsteps=45)
    ref, _ = grad_of_loss(g, sc, U0, dts, rho_t, every=20)
    for every in (1, 7, 45, 1000):
        got, _ = grad_of_loss(g, sc, U0, dts, rho_t, every=every)
        assert torch.equal(got, ref), every


def test_rollout_vjp_equals_autograd_path():
    say("rollout_vjp vs autograd 128^2, 40 steps")
    g, sc, U0, dts, rho_t = setup(128, 2, nsteps=40)
    ref, _ = grad_of_loss(g, sc, U0, dts, rho_t, every=20)
    with torch.no_grad():
        UT = AD.rollout(g, sc, U0, dts, every=20)
        lamT = torch.zeros_like(UT)
        lamT[0] = UT[0] - rho_t
        UT2, lam0 = AD.rollout_vjp(g, sc, U0, dts, lamT, every=20)
    assert torch.equal(UT2, UT)
    assert torch.equal(lam0, ref)


@pytest.mark.parametrize("order", [2, 1])
def test_gradient_matches_torch_autograd_128(order):
    """200 steps at 128^2, checkpoints every 20, against autograd of the restated stages."""
    say(f"rollout gradient vs torch autograd 128^2 order {order}, 200 steps")
    g, sc, U0, dts, rho_t = setup(128, order)
    got, loss = grad_of_loss(g, sc, U0, dts, rho_t, every=20)
    bc = tuple(g.bc)
    X = U0.clone().requires_grad_(True)
    U = X
    for dt in dts.tolist():
        U = TR.ssprk2_step(U, dt, g.nx, g.ny, g.dx, g.dy, g.gamma, bc, order)
    d = U[0] - rho_t
    ref_loss = 0.5 * (d * d).sum()
    ref_loss.backward()
    ref = X.grad.detach()
    say(f"  loss {loss:.15e} ref {float(ref_loss):.15e}")
    assert abs(loss - float(ref_loss)) <= 1e-10 * abs(float(ref_loss))
    err = (got - ref).abs().amax(dim=(1, 2)) / ref.abs().amax(dim=(1, 2))
    say(f"  per-field max err {err.tolist()}")
    assert float(err.max()) < 1e-9, err


def smooth_cells(n, seed, count, margin=0.17):
    """Cells at least `margin` away from the four-quadrant split lines (further than
    any wave travels in 200 CFL-0.4 steps)."""
    xs, ys, _ = PB.four_quadrant_params(seed)
    rng = np.random.default_rng(1)
    cells = []
    while len(cells) < count:
        i, j = (int(v) for v in rng.integers(2, n - 2, 2))
        if abs((i + 0.5) / n - xs) > margin and abs((j + 0.5) / n - ys) > margin:
            cells.append((j, i))
    return cells


def test_first_order_gradient_matches_fd_1024():
    say("rollout gradient vs central FD 1024^2 order 1, 200 steps, 64 cells")
    n, h = 1024, 1e-6
    g, sc, U0, dts, rho_t = setup(n, 1)
    grad, _ = grad_of_loss(g, sc, U0, dts, rho_t, every=20)
    grad = grad[0].cpu().numpy()
    errs = []
    for (j, i) in smooth_cells(n, 0, 64):
        with torch.no_grad():
            Up = U0.clone()
            Up[0, j, i] += h
            Um = U0.clone()
            Um[0, j, i] -= h
            rp = AD.rollout(g, sc, Up, dts, every=10 ** 6)[0]
            rm = AD.rollout(g, sc, Um, dts, every=10 ** 6)[0]
            # the loss difference formed cell by cell (no cancellation of two large sums)
            dl = float(((rp - rm) * (0.5 * (rp + rm) - rho_t)).sum())
        errs.append(abs(dl / (2 * h) - grad[j, i]) / abs(grad[j, i]))
    errs = np.array(errs)
    say(f"  FD rel err median {np.median(errs):.2e} max {errs.max():.2e}")
    assert errs.max() < 1e-6, errs


@pytest.mark.parametrize("base_amp", [1e-3, 0.2])
def test_second_order_gradient_matches_fd_1024(base_amp):
    """The benchmarked adjoint's own scheme (HLLC + minmod MUSCL, SSP-RK2, 200
    steps at 1024^2, checkpoints every 20): the directional derivative of the
    loss along four seeded smooth directions v over all four fields, central
    differences with h = 1e-6 against <grad, v>, 1e-6 relative. A +-h
    perturbation of one cell crosses minmod kinks somewhere in its cone of
    influence over 200 steps (in smooth regions the two minmod arguments
    differ by O(dx^2) ~ h), so single-cell differences mix branches at this
    resolution (first order: `test_first_order_gradient_matches_fd_1024`);
    along a smooth direction the crossings are a vanishing share of the
    response. Base states: the four-quadrant state with the smooth 1e-3
    perturbation the adjoint's benchmark uses, and a 0.2 one."""
    say(f"rollout directional derivative vs central FD 1024^2 order 2, 200 steps, base amp {base_amp}")
    n, h = 1024, 1e-6
    g, sc, U0, dts, rho_t = setup(n, 2, base_amp=base_amp)
    grad, _ = grad_of_loss(g, sc, U0, dts, rho_t, every=20)
    for k in range(4):
        v = to_dev(np.stack([PB.smooth_perturbation(g, 500 + 10 * k + f, 1.0) for f in range(4)]))
        v[1:3] *= 0.5
        adj = float((grad * v).sum())
        with torch.no_grad():
            rp = AD.rollout(g, sc, U0 + h * v, dts, every=10 ** 6)[0]
            rm = AD.rollout(g, sc, U0 - h * v, dts, every=10 ** 6)[0]
            fd = float(((rp - rm) * (0.5 * (rp + rm) - rho_t)).sum()) / (2 * h)
        say(f"  direction {k}: adjoint {adj:.12e} FD {fd:.12e} rel {abs(fd - adj) / abs(adj):.2e}")
        assert abs(fd - adj) <= 1e-6 * abs(adj)


def mhd_setup(n, metod, order, nsteps, bc=(api.BC_PERIODIC,) * 4):
    g = api.make_grid(n, n, bc=bc)
    sc = api.make_scheme(api.MHD, api.FLUX_HLLD if metod == 3 else api.FLUX_HLLC_MHD, order, api.SSPRK2, 0.4)
    U0 = to_dev(PB.orszag_tang(g, seed=1, amp=1e-3))
    Ut = to_dev(PB.orszag_tang(g, seed=2, amp=1e-2))
    dts, Usolver = AD.dt_schedule(g, sc, U0, nsteps)
    with torch.no_grad():
        rho_t = AD.rollout(g, sc, Ut, dts, every=nsteps + 1)[0].clone()
    return g, sc, U0, dts, rho_t, Usolver


def test_mhd_rollout_replays_solver_bitwise():
    say("mhd rollout vs native solver 64^2 HLLD, 30 steps")
    g, sc, U0, dts, _, Usolver = mhd_setup(64, 3, 2, 30, bc=(api.BC_REFLECTIVE, api.BC_TRANSMISSIVE,
                                                               api.BC_PERIODIC, api.BC_PERIODIC))
    assert torch.all(dts > 0)
    with torch.no_grad():
        UT = AD.rollout(g, sc, U0, dts, every=7)
    assert torch.equal(UT, Usolver)


@pytest.mark.parametrize("metod,order", [(3, 2), (2, 1)])
def test_mhd_gradient_matches_torch_autograd(metod, order):
    """Orszag-Tang 48^2, 24 SSP-RK2 steps, checkpoints every 5 (and every 1, 1000:
    bitwise the same), against autograd of the restated MHD stages."""
    say(f"mhd rollout gradient vs torch autograd 48^2 metod {metod} order {order}, 24 steps")
    g, sc, U0, dts, rho_t, _ = mhd_setup(48, metod, order, 24)
    got, loss = grad_of_loss(g, sc, U0, dts, rho_t, every=5)
    for every in (1, 1000):
        again, _ = grad_of_loss(g, sc, U0, dts, rho_t, every=every)
        assert torch.equal(again, got), every
    bc = tuple(g.bc)
    X = U0.cpu().clone().requires_grad_(True)
    U = X
    for dt in dts.tolist():
        U1 = TR.mhd_stage(U, None, 0.0, 1.0, dt, g.nx, g.ny, g.dx, g.dy, g.gamma, bc, metod, order)
        U = TR.mhd_stage(U1, U, 0.5, 0.5, dt, g.nx, g.ny, g.dx, g.dy, g.gamma, bc, metod, order)
    d = U[0] - rho_t.cpu()
    ref_loss = 0.5 * (d * d).sum()
    ref_loss.backward()
    ref = X.grad.detach()
    say(f"  loss {loss:.15e} ref {float(ref_loss):.15e}")
    assert abs(loss - float(ref_loss)) <= 1e-10 * abs(float(ref_loss))
    err = (got.cpu() - ref).abs().amax(dim=(1, 2)) / ref.abs().amax(dim=(1, 2)).clamp_min(1e-300)
    say(f"  per-field max err {err.tolist()}")
    assert float(err.max()) < 1e-9, err
