"""GPU parity of the MHD stage adjoint (gd2_stage_vjp with HLLC-MHD or HLLD).

* the GPU forward stage against the torch float64 restatement of the MHD
  stage (tests/torch_ref.py, itself pinned to the C oracle's stage and to
  central differences in tests/test_torch_ref_mhd.py) on the test states, so
  that autograd of the restatement linearises the map the GPU computes;
* J_L(U)^T V from the HIP kernels against torch float64 autograd of that
  restatement, for every boundary kind, both fluxes, both orders, random
  strong-field states (every HLLD region) and an Orszag-Tang state;
* the affine combination W = a*A + b*B + c*J^T V, with W aliasing A;
* the kernels' directional derivative against central differences of the
  GPU forward stage itself (transmissive and periodic sides: a reflective
  wall face sits on the flux discontinuity at SM = 0 when the normal field
  changes sign there, where only the one-sided derivative the forward takes
  is defined).
"""
import numpy as np
import pytest
import torch

from golikov_2d_amd import api
from golikov_2d_amd import problems as PB
from tests import torch_ref as TR
from tests.gpu_util import dev, say, to_dev, to_host

pytestmark = pytest.mark.gpu

T, P, R = api.BC_TRANSMISSIVE, api.BC_PERIODIC, api.BC_REFLECTIVE
FLUX = {2: api.FLUX_HLLC_MHD, 3: api.FLUX_HLLD}


def random_mhd(g, seed, bmax=2.5):
    rng = np.random.default_rng(seed)
    sh = (g.ny, g.nx)
    r, p = 1.0 + 0.3 * rng.random(sh), 1.0 + 0.3 * rng.random(sh)
    u, v = rng.uniform(-0.5, 0.5, sh), rng.uniform(-0.5, 0.5, sh)
    bx, by, bz = (rng.uniform(-bmax, bmax, sh) for _ in range(3))
    return np.ascontiguousarray(PB.mhd_from_prim(g.gamma, r, p, u, v, bx, by, bz))


def make_state(g, kind, seed):
    if kind == "ot":
        return PB.orszag_tang(g, seed=seed, amp=0.05)
    return random_mhd(g, seed)


def ref_vjp(g, bc, metod, order, U, V):
    Ut = torch.from_numpy(U).requires_grad_(True)
    L = TR.mhd_L(Ut, g.nx, g.ny, g.dx, g.dy, g.gamma, bc, metod, order)
    (W,) = torch.autograd.grad((L * torch.from_numpy(V)).sum(), Ut)
    return W.numpy()


def field_err(got, ref):
    return np.abs(got - ref).max(axis=(1, 2)) / np.maximum(np.abs(ref).max(axis=(1, 2)), 1e-300)


def scheme(metod, order):
    return api.make_scheme(api.MHD, FLUX[metod], order, api.SSPRK2, 0.4)


CASES = [
    (24, 19, (T, T, T, T), 3, 2, "random"),
    (24, 19, (P, P, P, P), 3, 2, "random"),
    (24, 19, (R, R, R, R), 3, 2, "random"),
    (21, 17, (R, T, T, R), 3, 2, "random"),
    (21, 17, (P, P, R, T), 3, 2, "random"),
    (23, 18, (T, R, R, T), 3, 2, "random"),
    (24, 19, (T, T, T, T), 2, 2, "random"),
    (24, 19, (R, R, R, R), 2, 2, "random"),
    (24, 19, (P, P, P, P), 3, 1, "random"),
    (24, 19, (R, T, R, T), 2, 1, "random"),
    (40, 33, (P, P, P, P), 3, 2, "ot"),
    (37, 30, (R, R, T, T), 3, 2, "ot"),
    (3, 5, (T, R, P, P), 3, 2, "random"),
    (2, 9, (R, R, T, R), 2, 2, "random"),
    # wider than one 60-column strip of the gather march: lane shifts across
    # strips, the ghost columns of the last strip (121: a one-column strip,
    # whose neighbour holds column nx on lane 63), periodic ghosts in other strips
    (121, 11, (P, P, T, T), 3, 2, "random"),
    (125, 13, (R, T, R, T), 2, 2, "random"),
    (61, 9, (T, T, P, P), 3, 1, "random"),
    (121, 7, (R, R, P, P), 3, 2, "random"),
]


@pytest.mark.parametrize("nx,ny,bc,metod,order,kind", CASES)
def test_mhd_stage_vjp_matches_autograd(nx, ny, bc, metod, order, kind):
    say(f"mhd vjp vs autograd {nx}x{ny} bc={bc} metod={metod} order={order} {kind}")
    g = api.make_grid(nx, ny, bc=bc)
    sc = scheme(metod, order)
    U = make_state(g, kind, seed=nx + ny)
    rng = np.random.default_rng(nx * ny)
    V = rng.normal(size=U.shape)
    # the GPU forward map is the restated one on this state
    dt = 1e-3
    out = torch.empty_like(to_dev(U))
    api.stage(g, sc, to_dev(U), None, out, 0.0, 1.0, dt)
    fwd = TR.mhd_stage(torch.from_numpy(U), None, 0.0, 1.0, dt, nx, ny, g.dx, g.dy, g.gamma, bc, metod,
                       order).numpy()
    assert field_err(to_host(out), fwd).max() < 1e-12
    ref = ref_vjp(g, bc, metod, order, U, V)
    W = torch.empty_like(to_dev(U))
    api.stage_vjp(g, sc, to_dev(U), to_dev(V), W, c=1.0)
    err = field_err(to_host(W), ref)
    assert err.max() < 1e-10, err


def test_mhd_stage_vjp_affine_combination_and_alias():
    say("mhd vjp affine combination")
    g = api.make_grid(30, 22, bc=(T, R, P, P))
    sc = scheme(3, 2)
    U = random_mhd(g, 3)
    rng = np.random.default_rng(5)
    V, A, B = (rng.normal(size=U.shape) for _ in range(3))
    a, b, c = 0.5, -1.25, 0.0375
    Ud, Vd = to_dev(U), to_dev(V)
    J = torch.empty_like(Ud)
    api.stage_vjp(g, sc, Ud, Vd, J, c=1.0)
    W = torch.empty_like(Ud)
    api.stage_vjp(g, sc, Ud, Vd, W, c=c, A=to_dev(A), a=a, B=to_dev(B), b=b)
    ref = a * A + b * B + c * to_host(J)
    assert np.abs(to_host(W) - ref).max() <= 1e-14 * np.abs(ref).max()
    Ad = to_dev(A)
    api.stage_vjp(g, sc, Ud, Vd, Ad, c=c, A=Ad, a=a, B=to_dev(B), b=b)  # W aliases A
    assert np.abs(to_host(Ad) - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [(30, 22), (70, 41), (3, 5)])
def test_mhd_stage_vjp_affine_term_that_is_v(shape):
    """An affine term given as V itself (the rollout's two stage adjoints:
    A = V, then B = V with W = A): W is bitwise that of the same call with a
    copy of V."""
    nx, ny = shape
    say(f"mhd vjp affine term = V {nx}x{ny}")
    g = api.make_grid(nx, ny, bc=(T, R, P, P))
    for metod, order in ((3, 2), (2, 2), (3, 1)):
        sc = scheme(metod, order)
        U = random_mhd(g, 7 + metod + order)
        rng = np.random.default_rng(11)
        V, A = (rng.normal(size=U.shape) for _ in range(2))
        Ud, Vd, Vc, Ad = to_dev(U), to_dev(V), to_dev(V), to_dev(A)
        W1, W2 = torch.empty_like(Ud), torch.empty_like(Ud)
        api.stage_vjp(g, sc, Ud, Vd, W1, c=0.3, A=Vd, a=0.5)
        api.stage_vjp(g, sc, Ud, Vd, W2, c=0.3, A=Vc, a=0.5)
        assert torch.equal(W1.view(torch.int64), W2.view(torch.int64)), (metod, order)
        A1, A2 = Ad.clone(), Ad.clone()
        api.stage_vjp(g, sc, Ud, Vd, A1, c=0.7, A=A1, a=0.5, B=Vd, b=1.0)
        api.stage_vjp(g, sc, Ud, Vd, A2, c=0.7, A=A2, a=0.5, B=Vc, b=1.0)
        assert torch.equal(A1.view(torch.int64), A2.view(torch.int64)), (metod, order)


@pytest.mark.parametrize("bc", [(T, T, T, T), (P, P, P, P), (P, P, T, T)])
@pytest.mark.parametrize("metod,order", [(3, 2), (2, 2), (3, 1)])
def test_mhd_stage_vjp_matches_gpu_forward_fd(bc, metod, order):
    """<W, E> with W = V + dt J_L^T V against central differences of
    <V, stage(U)> along E, stage(U) = U + dt L(U) on the GPU."""
    say(f"mhd vjp vs gpu fd bc={bc} metod={metod} order={order}")
    g = api.make_grid(33, 27, bc=bc)
    sc = scheme(metod, order)
    U = random_mhd(g, 11)
    rng = np.random.default_rng(12)
    V = rng.normal(size=U.shape)
    E = rng.normal(size=U.shape) * np.abs(U).mean(axis=(1, 2), keepdims=True)
    dt = 2e-3
    h = 1e-6

    def f(x):
        out = torch.empty_like(x)
        api.stage(g, sc, x, None, out, 0.0, 1.0, dt)
        return to_host(out)

    Ud, Ed = to_dev(U), to_dev(E)
    api.diagnostics(reset=True)
    fd = float((V * (f(Ud + h * Ed) - f(Ud - h * Ed))).sum() / (2 * h))
    d = api.diagnostics(reset=True)
    assert d["clamp_rho"] == 0 and d["clamp_p"] == 0
    W = torch.empty_like(Ud)
    api.stage_vjp(g, sc, Ud, to_dev(V), W, c=dt, A=to_dev(V), a=1.0)
    got = float((to_host(W) * E).sum())
    assert abs(got - fd) <= 1e-7 * abs(fd), (got, fd)


def test_mhd_stage_vjp_rejects_bad_arguments():
    g = api.make_grid(16, 16)
    sc = scheme(3, 2)
    U = to_dev(random_mhd(g, 1))
    with pytest.raises(api.GD2Error):
        api.stage_vjp(g, sc, U, U, U, c=1.0)  # W aliases U and V
    ws = torch.empty(8, dtype=torch.float64, device=dev())
    with pytest.raises(api.GD2Error):
        api.stage_vjp(g, sc, U, U.clone(), torch.empty_like(U), workspace=ws)  # workspace too small
