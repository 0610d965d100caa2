"""Pin the MHD part of the torch float64 restatement (tests/torch_ref.py),
the autograd reference of the MHD adjoint kernels:

* its stage values against the C oracle's MHD stage (o_stage: HLLDQ_Korolkov
  metod 2 / 3, minmod MUSCL, Powell term, mirrored wall faces) on the same
  inputs, for every boundary kind, both fluxes and both orders;
* its autograd vector-Jacobian products against central finite differences
  of its own forward map at cell-wise random states (minmod arguments away
  from ties and zeros).
"""
import numpy as np
import pytest
import torch

from golikov_2d_amd import problems as PB
from oracle import oracle as O
from tests import torch_ref as TR

T, P, R = O.BC_TRANSMISSIVE, O.BC_PERIODIC, O.BC_REFLECTIVE
BCS = [(T, T, T, T), (R, R, R, R), (P, P, P, P), (R, T, T, R), (P, P, R, T)]
FLUX = {2: O.FLUX_HLLC_MHD, 3: O.FLUX_HLLD}


def random_mhd(g, seed, bmax=2.5):
    """Cell-wise random primitives: rho, p in [1, 1.3], u, v in [-0.5, 0.5],
    Gaussian B components in [-bmax, bmax]."""
    rng = np.random.default_rng(seed)
    sh = (g.ny, g.nx)
    r, p = 1.0 + 0.3 * rng.random(sh), 1.0 + 0.3 * rng.random(sh)
    u, v = rng.uniform(-0.5, 0.5, sh), rng.uniform(-0.5, 0.5, sh)
    bx, by, bz = (rng.uniform(-bmax, bmax, sh) for _ in range(3))
    return np.ascontiguousarray(PB.mhd_from_prim(g.gamma, r, p, u, v, bx, by, bz))


def rel_err(got, ref):
    den = np.abs(ref).max(axis=(1, 2))
    return np.abs(got - ref).max(axis=(1, 2)) / np.where(den > 0, den, 1.0)


def stable_dt(g, U, cfl=0.3):
    r, u, v, p, bx, by, bz = (x.numpy() for x in TR.mhd_cons_to_prim(torch.from_numpy(U), g.gamma))
    cf = np.sqrt((g.gamma * p + (bx * bx + by * by + bz * bz) / TR.CPI4) / r)
    return cfl / np.max((np.abs(u) + cf) / g.dx + (np.abs(v) + cf) / g.dy)


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("metod", [2, 3])
@pytest.mark.parametrize("order", [1, 2])
def test_mhd_stage_matches_oracle(bc, metod, order):
    g = O.make_grid(19, 15, bc=bc)
    U = PB.orszag_tang(g, seed=3 + order, amp=0.05)
    Ub = random_mhd(g, seed=11)
    dt = stable_dt(g, U)
    sc = O.Scheme(O.PHYS_MHD, FLUX[metod], order, 0.4)
    for a, b, base in ((0.0, 1.0, None), (0.5, 0.5, Ub)):
        ref = O.stage(g, sc, U, base, a, b, dt)
        got = TR.mhd_stage(torch.from_numpy(U), None if base is None else torch.from_numpy(base), a, b, dt,
                           g.nx, g.ny, g.dx, g.dy, g.gamma, bc, metod, order).numpy()
        err = rel_err(got, ref)
        assert err.max() < 1e-12, err


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("metod", [2, 3])
def test_mhd_stage_matches_oracle_random_states(bc, metod):
    """Cell-wise random states with strong fields reach every HLLD region.

    At a reflective east / north wall the two face states are exact mirror
    images, so SM = 0 in exact arithmetic, and with the normal field changing
    sign across the wall the flux is discontinuous there (the tangential
    Maxwell stress -Bn Bt of the side taken). The restatement, like the native
    stage, forms the face values as q +- s/2 and gets SM = 0 exactly (the own
    side, as the reference's cell takes in exact arithmetic); the oracle's
    coordinate-based reconstruction (o_linear2) rounds the two states apart
    and may take either side. Those wall cells are compared when the flux is
    continuous (below: Bx = 0 on the wall column / By = 0 on the wall row)."""
    g = O.make_grid(17, 13, bc=bc)
    U = random_mhd(g, seed=7)
    dt = 0.2 * stable_dt(g, U)
    sc = O.Scheme(O.PHYS_MHD, FLUX[metod], 2, 0.4)
    ref = O.stage(g, sc, U, None, 0.0, 1.0, dt)
    got = TR.mhd_stage(torch.from_numpy(U), None, 0.0, 1.0, dt, g.nx, g.ny, g.dx, g.dy, g.gamma, bc, metod,
                       2).numpy()
    keep = np.ones((g.ny, g.nx), dtype=bool)
    if bc[1] == R:
        keep[:, -1] = False
    if bc[3] == R:
        keep[-1, :] = False
    den = np.abs(ref).max(axis=(1, 2))
    err = np.abs(got - ref)[:, keep].max(axis=1) / den
    assert err.max() < 1e-12, err
    if not keep.all():
        # no normal field on the reflective walls: the wall flux is continuous
        # at SM = 0 and the wall cells agree too
        U2 = U.copy()
        if bc[1] == R:
            U2[6, :, -1] -= U2[3, :, -1] ** 2 / TR.CPI8
            U2[3, :, -1] = 0.0
        if bc[3] == R:
            U2[6, -1, :] -= U2[4, -1, :] ** 2 / TR.CPI8
            U2[4, -1, :] = 0.0
        ref = O.stage(g, sc, U2, None, 0.0, 1.0, dt)
        got = TR.mhd_stage(torch.from_numpy(U2), None, 0.0, 1.0, dt, g.nx, g.ny, g.dx, g.dy, g.gamma, bc, metod,
                           2).numpy()
        assert rel_err(got, ref).max() < 1e-12


@pytest.mark.parametrize("bc", [(T, T, T, T), (P, P, P, P)])
@pytest.mark.parametrize("metod", [2, 3])
@pytest.mark.parametrize("order", [1, 2])
def test_mhd_autograd_vjp_matches_central_fd(bc, metod, order):
    g = O.make_grid(13, 11, bc=bc)
    U = random_mhd(g, seed=1)
    rng = np.random.default_rng(2)
    V = rng.normal(size=U.shape)
    E = rng.normal(size=U.shape) * np.abs(U).mean(axis=(1, 2), keepdims=True)
    h = 1e-7
    Ut = torch.from_numpy(U).requires_grad_(True)
    L = TR.mhd_L(Ut, g.nx, g.ny, g.dx, g.dy, g.gamma, bc, metod, order)
    (W,) = torch.autograd.grad((L * torch.from_numpy(V)).sum(), Ut)
    lhs = float((W.numpy() * E).sum())

    def Lv(x):
        return TR.mhd_L(torch.from_numpy(x), g.nx, g.ny, g.dx, g.dy, g.gamma, bc, metod, order).numpy()

    rhs = float((V * (Lv(U + h * E) - Lv(U - h * E))).sum() / (2 * h))
    assert abs(lhs - rhs) <= 1e-6 * abs(rhs), (lhs, rhs)
