"""GPU parity of the Euler stage adjoint (gd2_stage_vjp).

* J_L(U)^T V from the HIP kernels against torch float64 autograd of the
  restatement in tests/torch_ref.py (itself pinned to the C oracle's stage
  and to central differences in tests/test_torch_ref.py), for every boundary
  kind, both orders, several tilings, smooth and piecewise-constant states;
* the affine combination W = a*A + b*B + c*J^T V, with W aliasing A;
* the kernel's directional derivative against central differences of the
  GPU forward stage itself, at states away from minmod kinks;
* the SSP-RK2 step adjoint assembled from two stage VJPs against autograd
  through two restated stages.
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


def smooth_state(g, seed, amp=0.05):
    U = PB.four_quadrant(g, seed=seed).copy()
    d = PB.smooth_perturbation(g, seed + 17, amp)
    d2 = PB.smooth_perturbation(g, seed + 29, amp)
    U[0] *= 1.0 + d
    U[3] *= 1.0 + d2
    U[1] += 0.5 * d2 * U[0]
    U[2] -= 0.5 * d * U[0]
    return np.ascontiguousarray(U)


def noise_state(g, seed):
    rng = np.random.default_rng(seed)
    prim = [1.0 + 0.3 * rng.random((g.ny, g.nx)), 1.0 + 0.3 * rng.random((g.ny, g.nx)),
            rng.uniform(-0.5, 0.5, (g.ny, g.nx)), rng.uniform(-0.5, 0.5, (g.ny, g.nx))]
    return np.ascontiguousarray(PB.euler_from_prim(g.gamma, *prim))


def ref_vjp(g, bc, order, U, V):
    Ut = torch.from_numpy(U).requires_grad_(True)
    L = TR.euler_L(Ut, g.nx, g.ny, g.dx, g.dy, g.gamma, bc, order)
    (W,) = torch.autograd.grad((L * torch.from_numpy(V)).sum(), Ut)
    return W.numpy()


def field_err(got, ref):
    """per-field max |got - ref| / max |ref|"""
    return np.abs(got - ref).max(axis=(1, 2)) / np.maximum(np.abs(ref).max(axis=(1, 2)), 1e-300)


CASES = [
    (40, 33, (T, T, T, T), 2, "smooth"),
    (70, 45, (R, R, R, R), 2, "smooth"),
    (64, 48, (P, P, P, P), 2, "smooth"),
    (37, 129, (R, T, T, R), 2, "smooth"),
    (130, 70, (P, P, R, T), 2, "smooth"),
    (50, 30, (T, T, T, T), 1, "smooth"),
    (130, 70, (P, P, R, R), 1, "smooth"),
    (96, 80, (T, T, T, T), 2, "quadrants"),
    (61, 7, (R, R, T, T), 2, "smooth"),
    (3, 5, (T, R, P, P), 2, "smooth"),
    (128, 128, (T, T, T, T), 2, "smooth"),
    # strips of at least two rows at both ends: ghosts folded inside the kernel
    (70, 46, (R, R, R, R), 2, "smooth"),
    (37, 130, (R, T, T, R), 2, "smooth"),
    (62, 42, (T, R, R, T), 1, "quadrants"),
]


@pytest.mark.parametrize("nx,ny,bc,order,kind", CASES)
def test_stage_vjp_matches_autograd(nx, ny, bc, order, kind):
    say(f"stage_vjp vs autograd nx={nx} ny={ny} bc={bc} order={order} {kind}")
    g = api.make_grid(nx, ny, bc=bc)
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, order, api.SSPRK2, 0.4)
    U = smooth_state(g, seed=nx + ny) if kind == "smooth" else PB.four_quadrant(g, seed=4)
    V = np.random.default_rng(nx * ny).normal(size=U.shape)
    ref = ref_vjp(g, bc, order, U, V)
    Ud, Vd = to_dev(U), to_dev(V)
    W = torch.full_like(Ud, float("nan"))
    api.stage_vjp(g, sc, Ud, Vd, W)
    got = to_host(W)
    assert np.isfinite(got).all()
    err = field_err(got, ref)
    say(f"  max field err {err.max():.3e}")
    assert err.max() < 1e-11, err


def test_stage_vjp_affine_combination_and_alias():
    g = api.make_grid(90, 70, bc=(R, T, P, P))
    bc = tuple(g.bc)
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, 2, api.SSPRK2, 0.4)
    say("stage_vjp affine combination 90x70")
    U = smooth_state(g, seed=2)
    rng = np.random.default_rng(5)
    V, A, B = (rng.normal(size=U.shape) for _ in range(3))
    ref = 0.5 * A + 1.25 * B + 0.3 * ref_vjp(g, bc, 2, U, V)
    Ud, Vd, Ad, Bd = to_dev(U), to_dev(V), to_dev(A), to_dev(B)
    ws = api.vjp_workspace(g, dev())
    api.stage_vjp(g, sc, Ud, Vd, Ad, c=0.3, A=Ad, a=0.5, B=Bd, b=1.25, workspace=ws)  # W aliases A
    err = field_err(to_host(Ad), ref)
    assert err.max() < 1e-11, err
    # V may not alias W
    with pytest.raises(api.GD2Error):
        api.stage_vjp(g, sc, Ud, Vd, Vd, workspace=ws)


@pytest.mark.parametrize("bc,order", [((T, T, T, T), 2), ((P, P, R, R), 2), ((R, T, P, P), 1)])
def test_stage_vjp_matches_gpu_forward_fd(bc, order):
    """<J^T V, E> against <V, (L(U + hE) - L(U - hE)) / 2h>, L from the GPU forward stage."""
    g = api.make_grid(72, 66, bc=bc)
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, order, api.SSPRK2, 0.4)
    say(f"stage_vjp vs forward FD bc={bc} order={order}")
    U = noise_state(g, seed=9)
    rng = np.random.default_rng(10)
    V = rng.normal(size=U.shape)
    E = rng.normal(size=U.shape) * np.abs(U).mean(axis=(1, 2), keepdims=True)
    h, dt = 1e-7, 1e-4

    def Lgpu(x):
        xd = to_dev(x)
        out = torch.empty_like(xd)
        api.stage(g, sc, xd, None, out, 0.0, 1.0, dt=dt)
        return (to_host(out) - x) / dt

    rhs = float((V * (Lgpu(U + h * E) - Lgpu(U - h * E))).sum() / (2 * h))
    W = torch.empty_like(to_dev(U))
    api.stage_vjp(g, sc, to_dev(U), to_dev(V), W)
    lhs = float((to_host(W) * E).sum())
    say(f"  lhs {lhs:.12e} rhs {rhs:.12e}")
    assert abs(lhs - rhs) <= 1e-6 * abs(rhs), (lhs, rhs)


def test_ssprk2_step_adjoint_matches_autograd():
    g = api.make_grid(64, 64, bc=(T, T, T, T))
    bc = tuple(g.bc)
    sc = api.make_scheme(api.EULER, api.FLUX_HLLC, 2, api.SSPRK2, 0.4)
    say("ssprk2 step adjoint 64x64")
    U = smooth_state(g, seed=12)
    lam = np.random.default_rng(13).normal(size=U.shape)
    dt = 0.3 * g.dx / 3.0
    Ut = torch.from_numpy(U).requires_grad_(True)
    out = TR.ssprk2_step(Ut, dt, g.nx, g.ny, g.dx, g.dy, g.gamma, bc, 2)
    (ref,) = torch.autograd.grad((out * torch.from_numpy(lam)).sum(), Ut)
    ref = ref.numpy()

    Ud = to_dev(U)
    U1 = torch.empty_like(Ud)
    api.stage(g, sc, Ud, None, U1, 0.0, 1.0, dt=dt)
    ld = to_dev(lam)
    l1 = torch.empty_like(ld)
    ws = api.vjp_workspace(g, dev())
    api.stage_vjp(g, sc, U1, ld, l1, c=0.5 * dt, A=ld, a=0.5, workspace=ws)
    api.stage_vjp(g, sc, Ud, l1, ld, c=dt, A=ld, a=0.5, B=l1, b=1.0, workspace=ws)
    err = field_err(to_host(ld), ref)
    assert err.max() < 1e-11, err
