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


# ------------------------------------------------------------------ MHD
# Ideal MHD (vz = 0, Gaussian B, 7 conserved fields rho, rho u, rho v, Bx, By,
# Bz, E): minmod MUSCL on the primitives (r, u, v, p, bx, by, bz) with the
# rho/p fall-back, HLLDQ_Korolkov in metod 2 (HLLC-MHD) or 3 (HLLD)
# (Gas_Dinamic.cu:964-1510), the reflective west / south wall face evaluated
# mirrored, and the Powell source with the face-averaged neighbour field
# (kernel.cu:2559-2604).

CPI4 = 12.56637061435917295384
CPI8 = 25.13274122871834590768
SPI4 = 3.544907701811032


def mhd_cons_to_prim(U, gamma):
    r = U[0]
    u = U[1] / r
    v = U[2] / r
    bx, by, bz = U[3], U[4], U[5]
    p = (gamma - 1.0) * (U[6] - 0.5 * (U[1] * u + U[2] * v) - (bx * bx + by * by + bz * bz) / CPI8)
    return [r, u, v, p, bx, by, bz]


def _sel(c, a, b):
    return torch.where(c, a, b)


def hlld_face(metod, ro_L, p_L, u1, v1, w1, bn1, bt1, bm1, ro_R, p_R, u2, v2, w2, bn2, bt2, bm2, gamma):
    """Face-frame flux (mass, momentum n/t/m, energy, field n/t/m) of
    HLLDQ_Korolkov, metod 2 or 3, with the branch structure of
    Gas_Dinamic.cu:1250-1277 / 1434-1484 as elementwise selections."""
    inv_g1 = 1.0 / (gamma - 1.0)
    bb_L = bn1 * bn1 + bt1 * bt1 + bm1 * bm1
    bb_R = bn2 * bn2 + bt2 * bt2 + bm2 * bm2
    uu_L = 0.5 * (u1 * u1 + v1 * v1 + w1 * w1)
    uu_R = 0.5 * (u2 * u2 + v2 * v2 + w2 * w2)
    gpL, gpR = gamma * p_L + bb_L, gamma * p_R + bb_R
    cfL = torch.sqrt((gpL + torch.sqrt(gpL * gpL - 4.0 * gamma * p_L * (bn1 * bn1))) * (0.5 / ro_L))
    cfR = torch.sqrt((gpR + torch.sqrt(gpR * gpR - 4.0 * gamma * p_R * (bn2 * bn2))) * (0.5 / ro_R))
    cmax = _sel(cfL >= cfR, cfL, cfR)
    SL = _sel(u1 <= u2, u1, u2) - cmax
    SR = _sel(u1 >= u2, u1, u2) + cmax
    smax = _sel(SL.abs() >= SR.abs(), SL.abs(), SR.abs())
    pTL, pTR = p_L + 0.5 * bb_L, p_R + 0.5 * bb_R
    suR, suL = SR - u2, SL - u1
    mR, mL = suR * ro_R, suL * ro_L
    SM = (mR * u2 - mL * u1 - pTR + pTL) / (mR - mL)
    e1 = p_L * inv_g1 + ro_L * uu_L + 0.5 * bb_L
    e2 = p_R * inv_g1 + ro_R * uu_R + 0.5 * bb_R
    zero = torch.zeros_like(ro_L)
    FL = [ro_L * u1, ro_L * u1 * u1 + pTL - bn1 * bn1, ro_L * u1 * v1 - bn1 * bt1, ro_L * u1 * w1 - bn1 * bm1,
          (e1 + pTL) * u1 - bn1 * (u1 * bn1 + v1 * bt1 + w1 * bm1), zero, u1 * bt1 - v1 * bn1,
          u1 * bm1 - w1 * bn1]
    FR = [ro_R * u2, ro_R * u2 * u2 + pTR - bn2 * bn2, ro_R * u2 * v2 - bn2 * bt2, ro_R * u2 * w2 - bn2 * bm2,
          (e2 + pTR) * u2 - bn2 * (u2 * bn2 + v2 * bt2 + w2 * bm2), zero, u2 * bt2 - v2 * bn2,
          u2 * bm2 - w2 * bn2]
    UL = [ro_L, ro_L * u1, ro_L * v1, ro_L * w1, e1, bn1, bt1, bm1]
    UR = [ro_R, ro_R * u2, ro_R * v2, ro_R * w2, e2, bn2, bt2, bm2]
    iS = 1.0 / (SR - SL)
    bn = (SR * bn2 - SL * bn1) * iS
    bt = (SR * bt2 - SL * bt1 + FL[6] - FR[6]) * iS
    bm = (SR * bm2 - SL * bm1 + FL[7] - FR[7]) * iS
    bbn = bn * bn
    iLM, iRM = 1.0 / (SL - SM), 1.0 / (SR - SM)
    ro_LL = ro_L * suL * iLM
    ro_RR = ro_R * suR * iRM
    if metod == 2:
        sbv1 = u1 * bn1 + v1 * bt1 + w1 * bm1
        sbv2 = u2 * bn2 + v2 * bt2 + w2 * bm2
        UZ = [(SR * UR[i] - SL * UL[i] + FL[i] - FR[i]) * iS for i in range(4)]
        ptz = 0.5 * ((pTR + ro_R * suR * (SM - u2)) + (pTL + ro_L * suL * (SM - u1)))
        vv, ww = UZ[2] / UZ[0], UZ[3] / UZ[0]
        sbvz = (bn * UZ[1] + bt * UZ[2] + bm * UZ[3]) / UZ[0]
        ee2 = e2 * (suR * iRM) + (ptz * SM - pTR * u2 + bn * (sbv2 - sbvz)) * iRM
        ee1 = e1 * (suL * iLM) + (ptz * SM - pTL * u1 + bn * (sbv1 - sbvz)) * iLM
        left = (SL >= 0.0) | (SM >= 0.0)
        S = _sel(SL >= 0.0, zero, _sel(left, SL, _sel(SR > 0.0, SR, zero)))
        rs = _sel(left, ro_LL, ro_RR)
        es = _sel(left, ee1, ee2)
        Us = [rs, rs * SM, rs * vv, rs * ww, es, bn, bt, bm]
        P = [_sel(left, FL[i], FR[i]) + S * Us[i] - S * _sel(left, UL[i], UR[i]) for i in range(8)]
    else:
        PTT = (mR * pTL - mL * pTR + ro_L * ro_R * suR * suL * (u2 - u1)) / (mR - mL)
        ttL = ro_L * suL * (SL - SM) - bbn
        ttR = ro_R * suR * (SR - SM) - bbn
        okL, okR = ttL.abs() >= 0.00001, ttR.abs() >= 0.00001
        itL = 1.0 / _sel(okL, ttL, torch.ones_like(ttL))
        itR = 1.0 / _sel(okR, ttR, torch.ones_like(ttR))
        vLL = _sel(okL, v1 - bn * bt1 * (SM - u1) * itL, v1)
        wLL = _sel(okL, w1 - bn * bm1 * (SM - u1) * itL, w1)
        qL = (ro_L * suL * suL - bbn) * itL
        btt1, bmm1 = _sel(okL, bt1 * qL, zero), _sel(okL, bm1 * qL, zero)
        vRR = _sel(okR, v2 - bn * bt2 * (SM - u2) * itR, v2)
        wRR = _sel(okR, w2 - bn * bm2 * (SM - u2) * itR, w2)
        qR = (ro_R * suR * suR - bbn) * itR
        btt2, bmm2 = _sel(okR, bt2 * qR, zero), _sel(okR, bm2 * qR, zero)
        eLL = (e1 * suL + PTT * SM - pTL * u1 +
               bn * ((u1 * bn1 + v1 * bt1 + w1 * bm1) - (SM * bn + vLL * btt1 + wLL * bmm1))) * iLM
        eRR = (e2 * suR + PTT * SM - pTR * u2 +
               bn * ((u2 * bn2 + v2 * bt2 + w2 * bm2) - (SM * bn + vRR * btt2 + wRR * bmm2))) * iRM
        sLL, sRR = torch.sqrt(ro_LL), torch.sqrt(ro_RR)
        okB = bn.abs() > 0.000001
        idbn = torch.where(okB, torch.sign(bn.detach()), zero.detach())
        SLL = _sel(okB, SM - bn.abs() / sLL, SM)
        SRR = _sel(okB, SM + bn.abs() / sRR, SM)
        isum = 1.0 / (sLL + sRR)
        vLLL = (sLL * vLL + sRR * vRR + idbn * (btt2 - btt1)) * isum
        wLLL = (sLL * wLL + sRR * wRR + idbn * (bmm2 - bmm1)) * isum
        bttt = (sLL * btt2 + sRR * btt1 + idbn * sLL * sRR * (vRR - vLL)) * isum
        bmmm = (sLL * bmm2 + sRR * bmm1 + idbn * sLL * sRR * (wRR - wLL)) * isum
        vb3 = SM * bn + vLLL * bttt + wLLL * bmmm
        eLLL = eLL - idbn * sLL * ((SM * bn + vLL * btt1 + wLL * bmm1) - vb3)
        eRRR = eRR + idbn * sRR * ((SM * bn + vRR * btt2 + wRR * bmm2) - vb3)
        left = SM >= 0.0
        S1, S2 = _sel(left, SL, SR), _sel(left, SLL, SRR)
        rs = _sel(left, ro_LL, ro_RR)
        Us = [rs, rs * SM, rs * _sel(left, vLL, vRR), rs * _sel(left, wLL, wRR), _sel(left, eLL, eRR), bn,
              _sel(left, btt1, btt2), _sel(left, bmm1, bmm2)]
        Uss = [rs, rs * SM, rs * vLLL, rs * wLLL, _sel(left, eLLL, eRRR), bn, bttt, bmmm]
        outer = _sel(left, SL >= 0.0, ~((SRR > 0.0) | (SR > 0.0)))
        single = _sel(left, SLL >= 0.0, ~(SRR > 0.0))
        P = []
        for i in range(8):
            F, U = _sel(left, FL[i], FR[i]), _sel(left, UL[i], UR[i])
            P.append(_sel(outer, F, _sel(single, F + S1 * Us[i] - S1 * U,
                                         F + S2 * Uss[i] - (S2 - S1) * Us[i] - S1 * U)))
    P[5] = -smax * (bn2 - bn1)
    return P


def mhd_flux_lab(metod, xface, L, R, gamma):
    """Lab-frame fluxes (r, mx, my, bx, by, bz, e) of an x- or y-face from the
    face primitive lists L, R (r, u, v, p, bx, by, bz)."""
    s = 1.0 / SPI4
    z = torch.zeros_like(L[0])
    if xface:
        P = hlld_face(metod, L[0], L[3], L[1], L[2], z, L[4] * s, L[5] * s, L[6] * s,
                      R[0], R[3], R[1], R[2], z, R[4] * s, R[5] * s, R[6] * s, gamma)
        return [P[0], P[1], P[2], SPI4 * P[5], SPI4 * P[6], SPI4 * P[7], P[4]]
    P = hlld_face(metod, L[0], L[3], L[2], z, L[1], L[5] * s, L[6] * s, L[4] * s,
                  R[0], R[3], R[2], z, R[1], R[5] * s, R[6] * s, R[4] * s, gamma)
    return [P[0], P[3], P[1], SPI4 * P[7], SPI4 * P[5], SPI4 * P[6], P[4]]


def _wall_flux(metod, xface, L, R, gamma):
    """unmirror(flux(mirror(R), mirror(L))): a reflective west / south wall
    face evaluated from the interior cell's side."""
    nu, nb = (1, 4) if xface else (2, 5)
    mir = lambda q: [-a if f in (nu, nb) else a for f, a in enumerate(q)]
    F = mhd_flux_lab(metod, xface, mir(R), mir(L), gamma)
    odd = (0, 2, 4, 5, 6) if xface else (0, 1, 3, 5, 6)
    return [-a if f in odd else a for f, a in enumerate(F)]


def mhd_L(U, nx, ny, dx, dy, gamma, bc, metod=3, order=2):
    """L(U) = -(1/(dx dy)) (sum of outward face fluxes x lengths + Powell
    term) of the MHD stage; U: (7, ny, nx) float64."""
    q = mhd_cons_to_prim(U, gamma)
    dev = U.device
    ix, fx = ghost_index(nx, bc[0], bc[1])
    iy, fy = ghost_index(ny, bc[2], bc[3])
    ix_t, iy_t = torch.from_numpy(ix).to(dev), torch.from_numpy(iy).to(dev)
    sx = torch.from_numpy(np.where(fx, -1.0, 1.0)).to(dev)
    sy = torch.from_numpy(np.where(fy, -1.0, 1.0)).to(dev)

    qx = [a[:, ix_t] * sx if f in (1, 4) else a[:, ix_t] for f, a in enumerate(q)]
    Lp, Rm = _faces(qx, order)                      # (ny, nx+1): face k - 1/2
    F = mhd_flux_lab(metod, True, Lp, Rm, gamma)
    if bc[0] == BC_REFLECTIVE:
        W0 = _wall_flux(metod, True, [a[:, :1] for a in Lp], [a[:, :1] for a in Rm], gamma)
        F = [torch.cat([w, a[:, 1:]], dim=1) for w, a in zip(W0, F)]

    qy = [(a[iy_t, :] * sy[:, None]).T if f in (2, 5) else a[iy_t, :].T for f, a in enumerate(q)]
    Lp, Rm = _faces(qy, order)                      # (nx, ny+1)
    G = mhd_flux_lab(metod, False, Lp, Rm, gamma)
    if bc[2] == BC_REFLECTIVE:
        W0 = _wall_flux(metod, False, [a[:, :1] for a in Lp], [a[:, :1] for a in Rm], gamma)
        G = [torch.cat([w, a[:, 1:]], dim=1) for w, a in zip(W0, G)]
    G = [a.T for a in G]                            # (ny+1, nx)

    bx_ext = qx[4]                                   # columns -2 .. nx+1 (mirrored)
    by_ext = qy[5].T                                 # rows -2 .. ny+1
    pdiv = dy * (bx_ext[:, 3:nx + 3] - bx_ext[:, 1:nx + 1]) * 0.5 + dx * (by_ext[3:ny + 3, :] - by_ext[1:ny + 1, :]) * 0.5
    r, u, v, p, bx, by, bz = q
    zero = torch.zeros_like(r)
    K = [zero, bx / CPI4, by / CPI4, u, v, zero, (u * bx + v * by) / CPI4]
    out = []
    for f in range(7):
        dF = F[f][:, 1:] - F[f][:, :-1]
        dG = G[f][1:, :] - G[f][:-1, :]
        out.append(-(dF * dy + dG * dx + K[f] * pdiv) / (dx * dy))
    return torch.stack(out)


def mhd_stage(U, Ubase, a, b, dt, nx, ny, dx, dy, gamma, bc, metod=3, order=2):
    """a*Ubase + b*(U + dt L(U)) without the clamps."""
    W = U + dt * mhd_L(U, nx, ny, dx, dy, gamma, bc, metod, order)
    return b * W if Ubase is None else a * Ubase + b * W
