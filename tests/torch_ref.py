"""Differentiable float64 restatement of the Euler stage operator in torch.

Test infrastructure: the autograd reference for the adjoint kernels. It
restates the native stage (SoA conserved state U[4, ny, nx]; minmod MUSCL on
primitives with the rho/p <= 0 face fall-back, kernel.cu:3436-3544 and
kernel.cu:28; HLLC_Aleksashov with the averaged star pressure,
kernel.cu:853-1165; outward-flux-sum update, kernel.cu:1846-1874) with:

* minmod as alpha*dm + beta*dp with the branch weights recorded from the
  primal values (1/2-1/2 on |dm| == |dp|, 0-0 when the signs differ or an
  argument is zero), so autograd differentiates the branch the forward
  pass takes, with the same subgradient convention as the adjoint kernels;
* boundary ghosts as index gathers (clamp / mirror with a normal-velocity
  sign flip / wrap), so autograd folds ghost cotangents into their source
  cells.

The forward values are checked against the C oracle's stage on the CPU
(tests/test_torch_ref.py) before autograd through them is trusted.
"""
from __future__ import annotations

import numpy as np
import torch

BC_TRANSMISSIVE, BC_PERIODIC, BC_REFLECTIVE = 0, 1, 2


def ghost_index(n, lo, hi, width=2):
    """Source index and normal-flip flag of ext positions -width .. n+width-1."""
    ks = np.arange(-width, n + width)
    idx = np.empty(len(ks), dtype=np.int64)
    flip = np.zeros(len(ks), dtype=bool)
    for i, k in enumerate(ks):
        if k < 0:
            if lo == BC_PERIODIC:
                idx[i] = k % n
            elif lo == BC_REFLECTIVE:
                idx[i] = min(-1 - k, n - 1)
                flip[i] = True
            else:
                idx[i] = 0
        elif k >= n:
            if hi == BC_PERIODIC:
                idx[i] = k % n
            elif hi == BC_REFLECTIVE:
                idx[i] = max(2 * n - 1 - k, 0)
                flip[i] = True
            else:
                idx[i] = n - 1
        else:
            idx[i] = k
    return idx, flip


def cons_to_prim(U, gamma):
    r = U[0]
    u = U[1] / r
    v = U[2] / r
    p = (gamma - 1.0) * (U[3] - 0.5 * (U[1] * u + U[2] * v))
    return r, u, v, p


def minmod_recorded(dm, dp):
    with torch.no_grad():
        am, ap = dm.abs(), dp.abs()
        zero = ~(dm * dp > 0)
        one = torch.ones_like(dm)
        half = 0.5 * one
        nil = torch.zeros_like(dm)
        wa = torch.where(zero, nil, torch.where(am < ap, one, torch.where(am > ap, nil, half)))
        wb = torch.where(zero, nil, torch.where(ap < am, one, torch.where(ap > am, nil, half)))
    return wa * dm + wb * dp


def hllc(rL, unL, utL, pL, rR, unR, utR, pR, gamma):
    """HLLC flux in the face frame: (mass, normal mom., tangential mom., energy)."""
    inv_g1 = 1.0 / (gamma - 1.0)
    cL = torch.sqrt(gamma * pL / rL)
    cR = torch.sqrt(gamma * pR / rR)
    aL, aR = unL - cL, unR - cR
    bL, bR = unL + cL, unR + cR
    SL = torch.where(aL <= aR, aL, aR)
    SR = torch.where(bL >= bR, bL, bR)
    suL, suR = SL - unL, SR - unR
    mL, mR = suL * rL, suR * rR
    SM = (mR * unR - pR + pL - mL * unL) / (mR - mL)
    ptz = 0.5 * ((pR + mR * (SM - unR)) + (pL + mL * (SM - unL)))
    left = SM >= 0.0
    S = torch.where(left, SL, SR)
    r = torch.where(left, rL, rR)
    un = torch.where(left, unL, unR)
    ut = torch.where(left, utL, utR)
    p = torch.where(left, pL, pR)
    su = torch.where(left, suL, suR)
    sup = (SL > 0.0) | (SR < 0.0)
    e = p * inv_g1 + r * (0.5 * (un * un + ut * ut))
    den = torch.where(sup, torch.ones_like(S), S - SM)
    inv = 1.0 / den
    sm = su * inv
    rz = r * sm
    ez = e * sm + (ptz * SM - p * un) * inv
    Se = torch.where(sup, torch.zeros_like(S), S)
    run = r * un
    return (run + Se * (rz - r),
            run * un + p + Se * (SM * rz - run),
            run * ut + Se * (ut * rz - r * ut),
            (e + p) * un + Se * (ez - e))


def _faces(q, order):
    """q: 4 primitive tensors extended by 2 ghosts along the last axis
    (length n + 4). Returns (plus values of cells -1..n-1, minus values of
    cells 0..n), each a list of 4 tensors of length n + 1."""
    n = q[0].shape[-1] - 4
    plus, minus = [], []
    for f, a in enumerate(q):
        if order == 2:
            dm = a[..., 1:n + 3] - a[..., 0:n + 2]
            dp = a[..., 2:n + 4] - a[..., 1:n + 3]
            s = minmod_recorded(dm, dp)  # slopes of ext cells -1 .. n
        else:
            s = torch.zeros_like(a[..., 1:n + 3])
        qp = a[..., 1:n + 2] + 0.5 * s[..., 0:n + 1]
        qm = a[..., 2:n + 3] - 0.5 * s[..., 1:n + 2]
        if f in (0, 3):  # rho and p fall back to the cell value when not positive
            qp = torch.where(qp <= 0.0, a[..., 1:n + 2], qp)
            qm = torch.where(qm <= 0.0, a[..., 2:n + 3], qm)
        plus.append(qp)
        minus.append(qm)
    return plus, minus


def euler_L(U, nx, ny, dx, dy, gamma, bc, order=2):
    """Spatial operator L(U) = -(1/(dx dy)) sum of outward face fluxes x lengths.

    U: (4, ny, nx) float64 tensor. bc: (west, east, south, north) kinds.
    """
    r, u, v, p = cons_to_prim(U, gamma)
    dev = U.device
    ix, fx = ghost_index(nx, bc[0], bc[1])
    iy, fy = ghost_index(ny, bc[2], bc[3])
    ix_t = torch.from_numpy(ix).to(dev)
    iy_t = torch.from_numpy(iy).to(dev)
    sx = torch.from_numpy(np.where(fx, -1.0, 1.0)).to(dev)
    sy = torch.from_numpy(np.where(fy, -1.0, 1.0)).to(dev)

    # x faces: rows real, columns extended
    qx = [r[:, ix_t], u[:, ix_t] * sx, v[:, ix_t], p[:, ix_t]]
    Lp, Rm = _faces(qx, order)
    Fr, Fn, Ft, Fe = hllc(Lp[0], Lp[1], Lp[2], Lp[3], Rm[0], Rm[1], Rm[2], Rm[3], gamma)
    F = [Fr, Fn, Ft, Fe]  # (mass, x-mom, y-mom, energy), shape (ny, nx+1)

    # y faces: columns real, rows extended (work transposed: last axis = y)
    qy = [r[iy_t, :].T, u[iy_t, :].T, (v[iy_t, :] * sy[:, None]).T, p[iy_t, :].T]
    Lp, Rm = _faces(qy, order)
    Gr, Gn, Gt, Ge = hllc(Lp[0], Lp[2], Lp[1], Lp[3], Rm[0], Rm[2], Rm[1], Rm[3], gamma)
    G = [Gr.T, Gt.T, Gn.T, Ge.T]  # (mass, x-mom, y-mom, energy), shape (ny+1, nx)

    out = []
    for f in range(4):
        dF = F[f][:, 1:] - F[f][:, :-1]
        dG = G[f][1:, :] - G[f][:-1, :]
        out.append(-(dF * dy + dG * dx) / (dx * dy))
    return torch.stack(out)


def euler_stage(U, Ubase, a, b, dt, nx, ny, dx, dy, gamma, bc, order=2):
    """a*Ubase + b*(U + dt L(U)) without the clamps (they do not fire on the
    test states)."""
    W = U + dt * euler_L(U, nx, ny, dx, dy, gamma, bc, order)
    if Ubase is None:
        return b * W
    return a * Ubase + b * W


def ssprk2_step(U, dt, nx, ny, dx, dy, gamma, bc, order=2):
    U1 = euler_stage(U, None, 0.0, 1.0, dt, nx, ny, dx, dy, gamma, bc, order)
    return euler_stage(U1, U, 0.5, 0.5, dt, nx, ny, dx, dy, gamma, bc, order)
