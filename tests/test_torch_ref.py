"""Pin the torch float64 restatement (tests/torch_ref.py) used as the autograd
reference of the adjoint kernels.

* its stage values against the C oracle's stage (o_stage) on the same inputs
  for every boundary kind and both orders;
* its autograd vector-Jacobian products against central finite differences
  of its own forward map at states whose minmod arguments are all non-zero
  and unequal (no kinks within the step);
* the recorded-branch minmod against the oracle's minmod (values) and its
  branch weights against the textbook cases.
"""
import numpy as np
import pytest
import torch

from golikov_2d_amd import problems as PB
from oracle import oracle as O
from tests import torch_ref as TR

T, P, R = O.BC_TRANSMISSIVE, O.BC_PERIODIC, O.BC_REFLECTIVE
BCS = [(T, T, T, T), (R, R, R, R), (P, P, P, P), (R, T, T, R), (P, P, R, T)]


def state(g, seed, amp=0.05, quads=True):
    U = PB.four_quadrant(g, seed=seed) if quads else PB.uniform_euler(g, rho=1.0, p=1.0, u=0.3, v=-0.2)
    d = PB.smooth_perturbation(g, seed + 17, amp)
    d2 = PB.smooth_perturbation(g, seed + 29, amp)
    U = U.copy()
    U[0] *= 1.0 + d
    U[3] *= 1.0 + d2
    U[1] += 0.5 * d2 * U[0]
    U[2] -= 0.5 * d * U[0]
    return np.ascontiguousarray(U)


def stable_dt(g, U, cfl=0.3):
    r, u, v, p = (x.numpy() for x in TR.cons_to_prim(torch.from_numpy(U), g.gamma))
    c = np.sqrt(g.gamma * p / r)
    return cfl / np.max((np.abs(u) + c) / g.dx + (np.abs(v) + c) / g.dy)


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("order", [1, 2])
def test_stage_matches_oracle(bc, order):
    g = O.make_grid(23, 17, bc=bc)
    U = state(g, seed=3 + order)
    Ub = state(g, seed=11)
    dt = stable_dt(g, U)
    sc = O.Scheme(O.PHYS_EULER, O.FLUX_HLLC, order, 0.4)
    for a, b, base in ((0.0, 1.0, None), (0.5, 0.5, Ub)):
        ref = O.stage(g, sc, U, base, a, b, dt)
        got = TR.euler_stage(torch.from_numpy(U), None if base is None else torch.from_numpy(base), a, b, dt,
                             g.nx, g.ny, g.dx, g.dy, g.gamma, bc, order).numpy()
        err = np.abs(got - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))
        assert err.max() < 1e-13, err


def test_stage_matches_oracle_piecewise_constant():
    g = O.make_grid(40, 31, bc=(T, T, T, T))
    U = PB.four_quadrant(g, seed=5)
    dt = stable_dt(g, U)
    sc = O.Scheme(O.PHYS_EULER, O.FLUX_HLLC, 2, 0.4)
    ref = O.stage(g, sc, U, None, 0.0, 1.0, dt)
    got = TR.euler_stage(torch.from_numpy(U), None, 0.0, 1.0, dt, g.nx, g.ny, g.dx, g.dy, g.gamma,
                         (T, T, T, T), 2).numpy()
    err = np.abs(got - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))
    assert err.max() < 1e-13, err


def test_minmod_recorded_values_and_weights():
    rng = np.random.default_rng(0)
    a = rng.normal(size=4000)
    b = rng.normal(size=4000)
    a[:200] = 0.0
    b[200:400] = a[200:400]          # ties
    b[400:600] = -a[400:600]         # opposite signs, equal magnitude
    ta = torch.from_numpy(a).requires_grad_(True)
    tb = torch.from_numpy(b).requires_grad_(True)
    m = TR.minmod_recorded(ta, tb)
    ref = np.array([O.lib().o_minmod(x, y) for x, y in zip(a, b)])
    assert np.array_equal(m.detach().numpy(), ref)
    ga, gb = torch.autograd.grad(m.sum(), (ta, tb))
    ga, gb = ga.numpy(), gb.numpy()
    same = a * b > 0
    assert np.all(ga[~same] == 0) and np.all(gb[~same] == 0)
    tie = same & (np.abs(a) == np.abs(b))
    assert np.all(ga[tie] == 0.5) and np.all(gb[tie] == 0.5)
    lt = same & (np.abs(a) < np.abs(b))
    assert np.all(ga[lt] == 1) and np.all(gb[lt] == 0)
    gt = same & (np.abs(a) > np.abs(b))
    assert np.all(ga[gt] == 0) and np.all(gb[gt] == 1)


def _kink_free(U, g, bc, margin):
    """True when every minmod argument pair of U is away from a tie and from zero by margin."""
    r, u, v, p = TR.cons_to_prim(torch.from_numpy(U), g.gamma)
    ix, fx = TR.ghost_index(g.nx, bc[0], bc[1])
    iy, fy = TR.ghost_index(g.ny, bc[2], bc[3])
    # differences between a ghost and the cell it copies are zero whatever U is: not kinks
    sx = ~((ix[1:] == ix[:-1]) & (fx[1:] == fx[:-1]))
    sy = ~((iy[1:] == iy[:-1]) & (fy[1:] == fy[:-1]))
    worst = np.inf
    for f, q in enumerate((r, u, v, p)):
        q = q.numpy()
        qx = q[:, ix] * (np.where(fx, -1.0, 1.0) if f == 1 else 1.0)
        qy = q[iy, :] * (np.where(fy, -1.0, 1.0)[:, None] if f == 2 else 1.0)
        for d, live in ((np.diff(qx, axis=1), sx), (np.diff(qy, axis=0).T, sy)):
            dm, dp = d[:, :-1], d[:, 1:]
            both = live[:-1] & live[1:]
            worst = min(worst, np.abs(dm[:, live[:-1]]).min(), np.abs(dp[:, live[1:]]).min(),
                        np.abs(np.abs(dm) - np.abs(dp))[:, both].min())
    return worst > margin


@pytest.mark.parametrize("bc", [(T, T, T, T), (P, P, P, P)])
@pytest.mark.parametrize("order", [1, 2])
def test_autograd_vjp_matches_central_fd(bc, order):
    g = O.make_grid(14, 11, bc=bc)
    rng = np.random.default_rng(1)
    # cell-wise noise: minmod arguments are well separated from ties and zeros
    prim = [1.0 + 0.3 * rng.random((g.ny, g.nx)), 1.0 + 0.3 * rng.random((g.ny, g.nx)),
            rng.uniform(-0.5, 0.5, (g.ny, g.nx)), rng.uniform(-0.5, 0.5, (g.ny, g.nx))]
    U = np.ascontiguousarray(PB.euler_from_prim(g.gamma, *prim))
    V = rng.normal(size=U.shape)
    E = rng.normal(size=U.shape) * np.abs(U).mean(axis=(1, 2), keepdims=True)
    h = 1e-7
    if order == 2:
        # the check is only meaningful away from minmod kinks
        assert _kink_free(U, g, bc, margin=50 * h * np.abs(E).max())
    Ut = torch.from_numpy(U).requires_grad_(True)
    L = TR.euler_L(Ut, g.nx, g.ny, g.dx, g.dy, g.gamma, bc, order)
    (W,) = torch.autograd.grad((L * torch.from_numpy(V)).sum(), Ut)
    lhs = float((W.numpy() * E).sum())

    def Lv(x):
        return TR.euler_L(torch.from_numpy(x), g.nx, g.ny, g.dx, g.dy, g.gamma, bc, order).numpy()

    rhs = float((V * (Lv(U + h * E) - Lv(U - h * E))).sum() / (2 * h))
    assert abs(lhs - rhs) <= 1e-6 * abs(rhs), (lhs, rhs)
