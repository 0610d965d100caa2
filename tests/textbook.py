"""Independent textbook restatements used to pin the oracle.

These follow published methods, not the reference source:
  * exact Riemann solver: Toro, "Riemann Solvers and Numerical Methods for
    Fluid Dynamics", 3rd ed., ch. 4 (pressure function, Newton iteration,
    sampling);
  * HLLC: Toro ch. 10.4, star state eq. (10.39), wave speeds from the Davis
    estimate S_L = min(uL - cL, uR - cR), S_R = max(uL + cL, uR + cR);
  * HLLD: Miyoshi & Kusano, J. Comput. Phys. 208 (2005) 315, eqs. (38)-(63),
    Heaviside-Lorentz units;
  * HLLC for MHD: Li, J. Comput. Phys. 203 (2005) 344 ("HLLC-C"): HLLC
    with the contact speed S_M and star total pressure of the Euler HLLC,
    and the HLL averages for the star magnetic field, the star tangential
    velocities and the star (v.B);
  * both MHD solvers take the outer wave speeds S_L = min(uL, uR) -
    max(cfL, cfR), S_R = max(uL, uR) + max(cfL, cfR) (fast magnetosonic
    speeds). Where the two states' normal fields differ (a multi-dimensional
    scheme with cell-centred B), the star normal field is the HLL average
    (S_R BnR - S_L BnL) / (S_R - S_L), the value consistent with the integral
    form of the normal-field equation, whose physical flux is zero; the outer
    states keep their own normal field in their physical fluxes;
  * HLL: Harten, Lax & van Leer (SIAM Rev. 25, 1983) as in Toro eq. (10.21),
    with the bounds S_L = min(uL, uR) - max(cL, cR), S_R = max(uL, uR) +
    max(cL, cR) (Davis 1988, his simplest estimate).
All in pure Python floats (small inputs only).
"""
from __future__ import annotations

import math


# ------------------------------------------------------------------ Euler

def euler_phys_flux(gamma, rho, p, un, ut):
    E = p / (gamma - 1.0) + 0.5 * rho * (un * un + ut * ut)
    return (rho * un, rho * un * un + p, rho * un * ut, un * (E + p))


def hll(gamma, L, R):
    """L, R = (rho, p, un, ut). HLL flux (mass, normal mom., tangential mom., energy)."""
    rL, pL, uL, tL = L
    rR, pR, uR, tR = R
    cL, cR = math.sqrt(gamma * pL / rL), math.sqrt(gamma * pR / rR)
    SL = min(uL, uR) - max(cL, cR)
    SR = max(uL, uR) + max(cL, cR)
    FL, FR = euler_phys_flux(gamma, rL, pL, uL, tL), euler_phys_flux(gamma, rR, pR, uR, tR)
    if SL >= 0:
        return FL
    if SR <= 0:
        return FR
    UL = (rL, rL * uL, rL * tL, pL / (gamma - 1) + 0.5 * rL * (uL * uL + tL * tL))
    UR = (rR, rR * uR, rR * tR, pR / (gamma - 1) + 0.5 * rR * (uR * uR + tR * tR))
    return tuple((SR * fl - SL * fr + SL * SR * (ur - ul)) / (SR - SL) for fl, fr, ul, ur in zip(FL, FR, UL, UR))


def toro_exact(gamma, L, R, tol=1e-15):
    """L, R = (rho, p, u) along the normal. Returns (p*, u*)."""
    rL, pL, uL = L
    rR, pR, uR = R
    cL, cR = math.sqrt(gamma * pL / rL), math.sqrt(gamma * pR / rR)
    AL, AR = 2.0 / ((gamma + 1) * rL), 2.0 / ((gamma + 1) * rR)
    BL, BR = (gamma - 1) / (gamma + 1) * pL, (gamma - 1) / (gamma + 1) * pR

    def f(p, pk, rk, ck, Ak, Bk):
        if p > pk:
            q = math.sqrt(Ak / (p + Bk))
            return (p - pk) * q, q * (1 - (p - pk) / (2 * (Bk + p)))
        pr = p / pk
        return (2 * ck / (gamma - 1)) * (pr ** ((gamma - 1) / (2 * gamma)) - 1), \
            1 / (rk * ck) * pr ** (-(gamma + 1) / (2 * gamma))

    # two-rarefaction initial guess (Toro eq. 4.46)
    z = (gamma - 1) / (2 * gamma)
    p = ((cL + cR - 0.5 * (gamma - 1) * (uR - uL)) / (cL / pL ** z + cR / pR ** z)) ** (1 / z)
    p = max(p, 1e-14)
    for _ in range(200):
        fl, dl = f(p, pL, rL, cL, AL, BL)
        fr, dr = f(p, pR, rR, cR, AR, BR)
        pn = p - (fl + fr + uR - uL) / (dl + dr)
        if pn < 0:
            pn = tol
        if abs(pn - p) <= tol * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = f(p, pL, rL, cL, AL, BL)
    fr, _ = f(p, pR, rR, cR, AR, BR)
    return p, 0.5 * (uL + uR) + 0.5 * (fr - fl)


def toro_sample0(gamma, L, R):
    """State (rho, p, u) of the exact solution on x/t = 0 (Toro sec. 4.5)."""
    rL, pL, uL = L
    rR, pR, uR = R
    ps, us = toro_exact(gamma, L, R)
    g = gamma
    cL, cR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)
    S = 0.0
    if S <= us:  # left of contact
        if ps > pL:
            SL = uL - cL * math.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
            if S <= SL:
                return rL, pL, uL
            r = rL * (ps / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pL + 1)
            return r, ps, us
        SHL = uL - cL
        if S <= SHL:
            return rL, pL, uL
        csL = cL * (ps / pL) ** ((g - 1) / (2 * g))
        STL = us - csL
        if S > STL:
            return rL * (ps / pL) ** (1 / g), ps, us
        c = 2 / (g + 1) * (cL + (g - 1) / 2 * (uL - S))
        u = 2 / (g + 1) * (cL + (g - 1) / 2 * uL + S)
        r = rL * (c / cL) ** (2 / (g - 1))
        return r, pL * (c / cL) ** (2 * g / (g - 1)), u
    if ps > pR:
        SR = uR + cR * math.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
        if S >= SR:
            return rR, pR, uR
        r = rR * (ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1)
        return r, ps, us
    SHR = uR + cR
    if S >= SHR:
        return rR, pR, uR
    csR = cR * (ps / pR) ** ((g - 1) / (2 * g))
    STR = us + csR
    if S < STR:
        return rR * (ps / pR) ** (1 / g), ps, us
    c = 2 / (g + 1) * (cR - (g - 1) / 2 * (uR - S))
    u = 2 / (g + 1) * (-cR + (g - 1) / 2 * uR + S)
    r = rR * (c / cR) ** (2 / (g - 1))
    return r, pR * (c / cR) ** (2 * g / (g - 1)), u


def toro_hllc(gamma, L, R, speeds="davis", energy="rh", sm_left_u=None, pstar="side"):
    """L, R = (rho, p, un, ut). Returns normal-frame flux (F_rho, F_n, F_t, F_E).

    Toro's HLLC (ch. 10.4): contact speed S* (eq. 10.37), star states
    rho*_K = rho_K (S_K - u_K)/(S_K - S*), normal velocity S*, tangential
    velocity unchanged, flux F_K + S_K (U*_K - U_K) on either side of the
    contact. Options for the variants that differ from it:
      speeds = "davis": S_L = min(uL - cL, uR - cR), S_R = max(uL + cL, uR + cR);
               "simple": S_L = min(uL, uR) - max(cL, cR), S_R = max(uL, uR) +
               max(cL, cR) (Davis's simplest bounds, those of the HLL flux);
      energy = "rh": E*_K from the Rankine-Hugoniot condition across S_K,
               E*_K = ((S_K - u_K) E_K - p_K u_K + p*_K S*) / (S_K - S*) (eq.
               10.39 rearranged); "eos": E*_K = p*/(gamma-1) + rho*_K (S*^2 +
               ut_K^2)/2 with p* = pL + rhoL (S_L - uL)(S* - uL);
      pstar = "side": p*_K = p_K + rho_K (S_K - u_K)(S* - u_K), the two being
               equal when S* is eq. 10.37; "mean": both sides take the mean of
               the two one-sided values;
      sm_left_u: replaces uL in the left momentum term rho_L uL (S_L - uL) of
               the numerator of S* (a variant that puts another velocity there).
    """
    rL, pL, uL, vL = L
    rR, pR, uR, vR = R
    cL, cR = math.sqrt(gamma * pL / rL), math.sqrt(gamma * pR / rR)
    if speeds == "davis":
        SL, SR = min(uL - cL, uR - cR), max(uL + cL, uR + cR)
    else:
        SL, SR = min(uL, uR) - max(cL, cR), max(uL, uR) + max(cL, cR)
    uLm = uL if sm_left_u is None else sm_left_u
    Ss = (pR - pL + rL * uLm * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR))
    FL = euler_phys_flux(gamma, rL, pL, uL, vL)
    FR = euler_phys_flux(gamma, rR, pR, uR, vR)
    ps = pL + rL * (SL - uL) * (Ss - uL)
    psR = pR + rR * (SR - uR) * (Ss - uR)
    pmean = 0.5 * (ps + psR)

    def ustar(r, p, u, v, S):
        E = p / (gamma - 1) + 0.5 * r * (u * u + v * v)
        f = r * (S - u) / (S - Ss)
        if energy == "rh":
            pk = pmean if pstar == "mean" else p + r * (S - u) * (Ss - u)
            Es = ((S - u) * E - p * u + pk * Ss) / (S - Ss)
        else:
            Es = ps / (gamma - 1) + 0.5 * f * (Ss * Ss + v * v)
        return (f, f * Ss, f * v, Es)

    def cons(r, p, u, v):
        return (r, r * u, r * v, p / (gamma - 1) + 0.5 * r * (u * u + v * v))

    if SL >= 0:
        F = FL
    elif Ss >= 0:
        Us, U = ustar(rL, pL, uL, vL, SL), cons(rL, pL, uL, vL)
        F = tuple(FL[i] + SL * (Us[i] - U[i]) for i in range(4))
    elif SR >= 0:
        Us, U = ustar(rR, pR, uR, vR, SR), cons(rR, pR, uR, vR)
        F = tuple(FR[i] + SR * (Us[i] - U[i]) for i in range(4))
    else:
        F = FR
    return (F[0], F[1], F[2], F[3])


def hllc_contact_side(gamma, L, R):
    """+1 when x/t = 0 lies left of the contact of Toro's HLLC (Davis speeds), -1 right."""
    rL, pL, uL, _ = L
    rR, pR, uR, _ = R
    cL, cR = math.sqrt(gamma * pL / rL), math.sqrt(gamma * pR / rR)
    SL, SR = min(uL - cL, uR - cR), max(uL + cL, uR + cR)
    Ss = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR))
    return 1 if Ss > 0 else -1


# ------------------------------------------------------------------ MHD

def _mhd_side(g, Q):
    """Face-frame state Q = (rho, p, un, vt, wm, Bn, Bt, Bm), Heaviside-Lorentz B:
    its conserved vector and physical normal flux over the 8 components
    (rho, mn, mt, mm, Bn, Bt, Bm, E) and the derived scalars."""
    r, p, u, v, w, bn, bt, bm = Q
    b2 = bn * bn + bt * bt + bm * bm
    pt = p + 0.5 * b2
    E = p / (g - 1) + 0.5 * r * (u * u + v * v + w * w) + 0.5 * b2
    cf = math.sqrt((g * p + b2 + math.sqrt((g * p + b2) ** 2 - 4 * g * p * bn * bn)) / (2 * r))
    vb = u * bn + v * bt + w * bm
    U = (r, r * u, r * v, r * w, bn, bt, bm, E)
    F = (r * u, r * u * u + pt - bn * bn, r * u * v - bn * bt, r * u * w - bn * bm,
         0.0, u * bt - v * bn, u * bm - w * bn, (E + pt) * u - bn * vb)
    return dict(r=r, p=p, u=u, v=v, w=w, bn=bn, bt=bt, bm=bm, pt=pt, E=E, cf=cf, vb=vb, U=U, F=F)


def _outer_speeds(A, B):
    cf = max(A["cf"], B["cf"])
    return min(A["u"], B["u"]) - cf, max(A["u"], B["u"]) + cf


def _hll_state(A, B, SL, SR):
    return tuple((SR * ur - SL * ul + fl - fr) / (SR - SL) for ul, ur, fl, fr in zip(A["U"], B["U"], A["F"], B["F"]))


def _contact_speed(A, B, SL, SR):
    sL, sR = (SL - A["u"]) * A["r"], (SR - B["u"]) * B["r"]
    return (sR * B["u"] - sL * A["u"] - B["pt"] + A["pt"]) / (sR - sL)


def li_hllc(gamma, L, R):
    """Li (2005) HLLC-C for MHD in the face frame, Heaviside-Lorentz B.

    L, R = (rho, p, un, vt, wm, Bn, Bt, Bm). Returns the 8 normal-frame fluxes
    (rho, mn, mt, mm, Bn, Bt, Bm, E), with the physical (zero) normal-field
    flux, and (S_L, S_R).

    Star states, K = L, R:
      rho*_K = rho_K (S_K - u_K) / (S_K - S_M), normal velocity S_M;
      B* = B_hll (all three components);
      tangential velocities v*, w* = (rho v)_hll / rho_hll, (rho w)_hll / rho_hll;
      p*_T = p_TK + rho_K (S_K - u_K)(S_M - u_K) (the same from both sides
        when the normal field is continuous; their mean otherwise);
      E*_K = [(S_K - u_K) E_K - p_TK u_K + p*_T S_M
              + B*_n ((v.B)_K - (v.B)_hll)] / (S_K - S_M),
        (v.B)_hll = v_hll . B_hll with v_hll = (rho v)_hll / rho_hll;
    and the flux F_K + S_K (U*_K - U_K) on either side of the contact.
    """
    A, B = _mhd_side(gamma, L), _mhd_side(gamma, R)
    SL, SR = _outer_speeds(A, B)
    Uh = _hll_state(A, B, SL, SR)
    SM = _contact_speed(A, B, SL, SR)
    vh = (Uh[1] / Uh[0], Uh[2] / Uh[0], Uh[3] / Uh[0])
    bh = (Uh[4], Uh[5], Uh[6])
    vb_h = vh[0] * bh[0] + vh[1] * bh[1] + vh[2] * bh[2]
    pts = 0.5 * ((A["pt"] + A["r"] * (SL - A["u"]) * (SM - A["u"])) +
                 (B["pt"] + B["r"] * (SR - B["u"]) * (SM - B["u"])))

    def star(K, S):
        rs = K["r"] * (S - K["u"]) / (S - SM)
        Es = ((S - K["u"]) * K["E"] - K["pt"] * K["u"] + pts * SM + bh[0] * (K["vb"] - vb_h)) / (S - SM)
        return (rs, rs * SM, rs * vh[1], rs * vh[2], bh[0], bh[1], bh[2], Es)

    if SL >= 0:
        F = A["F"]
    elif SM >= 0:
        Us = star(A, SL)
        F = tuple(A["F"][k] + SL * (Us[k] - A["U"][k]) for k in range(8))
    elif SR > 0:
        Us = star(B, SR)
        F = tuple(B["F"][k] + SR * (Us[k] - B["U"][k]) for k in range(8))
    else:
        F = B["F"]
    return F, (SL, SR)


def mk_hlld(gamma, L, R, deg_abs=None, alfven_abs=0.0):
    """Miyoshi-Kusano HLLD, Heaviside-Lorentz B.

    L, R = (rho, p, un, vt, wm, Bn, Bt, Bm) in the face frame. Returns the 7
    fluxes (rho, mn, mt, mm, Bt, Bm, E) and (SL, SR). The star normal field
    B_x is the HLL average of the two sides' normal fields (their common
    value when they agree).

    Degenerate cases (M-K sec. 4.3): where rho_K (S_K - u_K)(S_K - S_M) - B_x^2
    vanishes, v*_K = v_K and B*_tK = 0; with deg_abs given, "vanishes" means
    |.| < deg_abs, else |.| below 1e-12 of its first term. Where B_x vanishes
    the Alfven waves merge with the contact (S*_L = S*_R = S_M); "vanishes"
    means |B_x| <= alfven_abs.
    """
    g = gamma
    A, B = _mhd_side(g, L), _mhd_side(g, R)
    SL, SR = _outer_speeds(A, B)
    SM = _contact_speed(A, B, SL, SR)
    Bx = (SR * B["bn"] - SL * A["bn"]) / (SR - SL)
    rL, uL, rR, uR = A["r"], A["u"], B["r"], B["u"]
    pts = ((SR - uR) * rR * A["pt"] - (SL - uL) * rL * B["pt"] + rL * rR * (SR - uR) * (SL - uL) * (uR - uL)) / \
          ((SR - uR) * rR - (SL - uL) * rL)

    def star(K, S):
        r, u, v, w, bt, bm = K["r"], K["u"], K["v"], K["w"], K["bt"], K["bm"]
        rs = r * (S - u) / (S - SM)
        den = r * (S - u) * (S - SM) - Bx * Bx
        tol = deg_abs if deg_abs is not None else 1e-12 * max(1.0, abs(r * (S - u) * (S - SM)))
        if abs(den) < tol:
            vs, ws, bts, bms = v, w, 0.0, 0.0
        else:
            vs = v - Bx * bt * (SM - u) / den
            ws = w - Bx * bm * (SM - u) / den
            bts = bt * (r * (S - u) ** 2 - Bx * Bx) / den
            bms = bm * (r * (S - u) ** 2 - Bx * Bx) / den
        vbs = SM * Bx + vs * bts + ws * bms
        Es = ((S - u) * K["E"] - K["pt"] * u + pts * SM + Bx * (K["vb"] - vbs)) / (S - SM)
        return rs, vs, ws, bts, bms, Es, (rs, rs * SM, rs * vs, rs * ws, bts, bms, Es)

    def seven(t8):  # drop the normal-field component
        return (t8[0], t8[1], t8[2], t8[3], t8[5], t8[6], t8[7])

    FL, UL, FR, UR = seven(A["F"]), seven(A["U"]), seven(B["F"]), seven(B["U"])
    rsL, vsL, wsL, btsL, bmsL, EsL, UsL = star(A, SL)
    rsR, vsR, wsR, btsR, bmsR, EsR, UsR = star(B, SR)
    sL, sR = math.sqrt(rsL), math.sqrt(rsR)
    SsL = SM - abs(Bx) / sL
    SsR = SM + abs(Bx) / sR
    if abs(Bx) <= alfven_abs:
        SsL = SsR = SM
        sgn = 0.0
    else:
        sgn = math.copysign(1.0, Bx)
    vss = (sL * vsL + sR * vsR + (btsR - btsL) * sgn) / (sL + sR)
    wss = (sL * wsL + sR * wsR + (bmsR - bmsL) * sgn) / (sL + sR)
    btss = (sL * btsR + sR * btsL + sL * sR * (vsR - vsL) * sgn) / (sL + sR)
    bmss = (sL * bmsR + sR * bmsL + sL * sR * (wsR - wsL) * sgn) / (sL + sR)
    vbss = SM * Bx + vss * btss + wss * bmss
    EssL = EsL - sL * (SM * Bx + vsL * btsL + wsL * bmsL - vbss) * sgn
    EssR = EsR + sR * (SM * Bx + vsR * btsR + wsR * bmsR - vbss) * sgn
    UssL = (rsL, rsL * SM, rsL * vss, rsL * wss, btss, bmss, EssL)
    UssR = (rsR, rsR * SM, rsR * vss, rsR * wss, btss, bmss, EssR)
    if SL >= 0:
        F = FL
    elif SsL >= 0:
        F = tuple(FL[i] + SL * (UsL[i] - UL[i]) for i in range(7))
    elif SM >= 0:
        F = tuple(FL[i] + SsL * UssL[i] - (SsL - SL) * UsL[i] - SL * UL[i] for i in range(7))
    elif SsR > 0:
        F = tuple(FR[i] + SsR * UssR[i] - (SsR - SR) * UsR[i] - SR * UR[i] for i in range(7))
    elif SR > 0:
        F = tuple(FR[i] + SR * (UsR[i] - UR[i]) for i in range(7))
    else:
        F = FR
    return F, (SL, SR)


def mhd_phys_flux_hl(gamma, r, p, u, v, w, bn, bt, bm):
    b2 = bn * bn + bt * bt + bm * bm
    pt = p + 0.5 * b2
    E = p / (gamma - 1) + 0.5 * r * (u * u + v * v + w * w) + 0.5 * b2
    vb = u * bn + v * bt + w * bm
    return (r * u, r * u * u + pt - bn * bn, r * u * v - bn * bt, r * u * w - bn * bm,
            u * bt - v * bn, u * bm - w * bn, (E + pt) * u - bn * vb)


# ------------------------------------------------------------------ Sod

def sod_exact(gamma, x, t, x0=0.5, L=(1.0, 1.0, 0.0), R=(0.125, 0.1, 0.0)):
    """Exact Sod profile (rho, p, u) at points x, time t (Toro sampling)."""
    out = []
    for xi in x:
        S = (xi - x0) / t
        # shift the frame: sample at speed S by Galilean transform
        Ls = (L[0], L[1], L[2] - S)
        Rs = (R[0], R[1], R[2] - S)
        r, p, u = toro_sample0(gamma, Ls, Rs)
        out.append((r, p, u + S))
    return out
