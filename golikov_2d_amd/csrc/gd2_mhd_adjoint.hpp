// gd2_mhd_adjoint.hpp -- discrete adjoint (vector-Jacobian product) of the
// ideal-MHD stage operator of gd2_mhd.hpp (HLLC-MHD or HLLD face flux,
// minmod MUSCL on the 7 primitives, Powell source) on gfx950.
//
// For L(U) = -(1/dV) (sum of face fluxes x face lengths + Powell term) the
// kernels compute W = a*A + b*B + c * J_L(U)^T V in two passes:
//
//  1. a face march (one wave per strip of columns, up its rows; each lane
//     reverses one y-face and one x-face per row): the face cotangent
//     phi = (h/dV)(V_right - V_left) (zero on the side of a ghost cell) is
//     taken through the face flux to its two reconstructed states by one
//     reverse sweep of hlld_face (hand-reversed, the same branches as the
//     forward); each product phi . dF/dx_k goes to plane k of a face buffer
//     (planes 0..6: left state, 7..13: right).
//     Reflective west / south wall faces are evaluated mirrored, as the
//     forward stage does.
//  2. a marching gather (one wave per strip of columns, up its rows): every
//     cell spreads the cotangents of its two face states on each line over
//     its minmod stencil once (the branch the forward takes, 1/2-1/2 on ties,
//     the rho / p <= 0 face fall-back as a copy); a cell's gather is its own
//     share plus its neighbours' shares (lane shifts along x, a register
//     march along y); it adds the derivatives of the Powell term, maps
//     primitive to conserved cotangents and forms W. The gathers of the
//     boundary ghost cells go to side buffers, and a fold kernel adds them to
//     the cells they copy (mirrored components negated), with the Powell
//     terms of the edge cells that read those copies.
//
// The forward stage's clamps are not differentiated (they do not fire at the
// smooth states an adjoint is linearised at).
#pragma once

#include "gd2_adjoint.hpp"
#include "gd2_mhd.hpp"

namespace gd2 {

// The forward quantities of hlld_face_vjp, from the 16 face-frame inputs IN
// (hlld_face's expressions); expanded once for the flux-specific part of the
// reverse sweep and once for the common part.
#define GD2_HLLD_VJP_FORWARD(IN) \
    [[maybe_unused]] const double ro_L = IN[0], p_L = IN[1], u1 = IN[2], v1 = IN[3], w1 = IN[4], bn1 = IN[5], bt1 = IN[6], \
                 bm1 = IN[7]; \
    [[maybe_unused]] const double ro_R = IN[8], p_R = IN[9], u2 = IN[10], v2 = IN[11], w2 = IN[12], bn2 = IN[13], bt2 = IN[14], \
                 bm2 = IN[15]; \
    /* ---------------------------------------------------------------- forward */ \
    [[maybe_unused]] const double bb_L = bn1 * bn1 + bt1 * bt1 + bm1 * bm1; \
    [[maybe_unused]] const double bb_R = bn2 * bn2 + bt2 * bt2 + bm2 * bm2; \
    [[maybe_unused]] const double uu_L = 0.5 * (u1 * u1 + v1 * v1 + w1 * w1); \
    [[maybe_unused]] const double uu_R = 0.5 * (u2 * u2 + v2 * v2 + w2 * w2); \
    [[maybe_unused]] const double gpL = gamma * p_L + bb_L, gpR = gamma * p_R + bb_R; \
    [[maybe_unused]] const double iroL = fast_rcp(ro_L), iroR = fast_rcp(ro_R); \
    [[maybe_unused]] const double g4 = 4.0 * gamma; \
    [[maybe_unused]] const double dL = gpL * gpL - g4 * p_L * (bn1 * bn1), dR = gpR * gpR - g4 * p_R * (bn2 * bn2); \
    [[maybe_unused]] const double sdL = fast_sqrt(dL), sdR = fast_sqrt(dR); \
    [[maybe_unused]] const double aL = (gpL + sdL) * (0.5 * iroL), aR = (gpR + sdR) * (0.5 * iroR); \
    [[maybe_unused]] const double cfL = fast_sqrt(aL), cfR = fast_sqrt(aR); \
    [[maybe_unused]] const bool cL = cfL >= cfR; /* tmax picks its first argument on ties */ \
    [[maybe_unused]] const double cmax = cL ? cfL : cfR; \
    [[maybe_unused]] const bool uminL = u1 <= u2, umaxL = u1 >= u2; \
    [[maybe_unused]] const double SL = (uminL ? u1 : u2) - cmax; \
    [[maybe_unused]] const double SR = (umaxL ? u1 : u2) + cmax; \
    [[maybe_unused]] const double aSL = SL < 0.0 ? -SL : SL, aSR = SR < 0.0 ? -SR : SR; \
    [[maybe_unused]] const bool smL = aSL >= aSR; \
    [[maybe_unused]] const double smax = smL ? aSL : aSR; \
    [[maybe_unused]] const double pTL = p_L + 0.5 * bb_L, pTR = p_R + 0.5 * bb_R; \
    [[maybe_unused]] const double suR = SR - u2, suL = SL - u1; \
    [[maybe_unused]] const double mR = suR * ro_R, mL = suL * ro_L; \
    [[maybe_unused]] const double Dm = mR - mL; \
    [[maybe_unused]] const double idm = fast_rcp(Dm); \
    [[maybe_unused]] const double smR = mR * u2, smLv = mL * u1; \
    [[maybe_unused]] const double Nn = smR - smLv - pTR + pTL; \
    [[maybe_unused]] const double SM = Nn * idm; \
    [[maybe_unused]] const double e1 = p_L * inv_g1 + ro_L * uu_L + 0.5 * bb_L; \
    [[maybe_unused]] const double e2 = p_R * inv_g1 + ro_R * uu_R + 0.5 * bb_R; \
    [[maybe_unused]] double FL[8], FR[8], UL[8], UR[8]; \
    FL[0] = ro_L * u1; \
    FL[1] = ro_L * u1 * u1 + pTL - bn1 * bn1; \
    FL[2] = ro_L * u1 * v1 - bn1 * bt1; \
    FL[3] = ro_L * u1 * w1 - bn1 * bm1; \
    [[maybe_unused]] const double sbv1 = u1 * bn1 + v1 * bt1 + w1 * bm1; \
    FL[4] = (e1 + pTL) * u1 - bn1 * sbv1; \
    FL[5] = 0.0; \
    FL[6] = u1 * bt1 - v1 * bn1; \
    FL[7] = u1 * bm1 - w1 * bn1; \
    FR[0] = ro_R * u2; \
    FR[1] = ro_R * u2 * u2 + pTR - bn2 * bn2; \
    FR[2] = ro_R * u2 * v2 - bn2 * bt2; \
    FR[3] = ro_R * u2 * w2 - bn2 * bm2; \
    [[maybe_unused]] const double sbv2 = u2 * bn2 + v2 * bt2 + w2 * bm2; \
    FR[4] = (e2 + pTR) * u2 - bn2 * sbv2; \
    FR[5] = 0.0; \
    FR[6] = u2 * bt2 - v2 * bn2; \
    FR[7] = u2 * bm2 - w2 * bn2; \
    UL[0] = ro_L; UL[1] = ro_L * u1; UL[2] = ro_L * v1; UL[3] = ro_L * w1; \
    UL[4] = e1; UL[5] = bn1; UL[6] = bt1; UL[7] = bm1; \
    UR[0] = ro_R; UR[1] = ro_R * u2; UR[2] = ro_R * v2; UR[3] = ro_R * w2; \
    UR[4] = e2; UR[5] = bn2; UR[6] = bt2; UR[7] = bm2; \
    [[maybe_unused]] const double DS = SR - SL; \
    [[maybe_unused]] const double iS = fast_rcp(DS); \
    [[maybe_unused]] const double nbn = SR * bn2 - SL * bn1; \
    [[maybe_unused]] const double nbt = SR * bt2 - SL * bt1 + FL[6] - FR[6]; \
    [[maybe_unused]] const double nbm = SR * bm2 - SL * bm1 + FL[7] - FR[7]; \
    [[maybe_unused]] const double bn = nbn * iS, bt = nbt * iS, bm = nbm * iS; \
    [[maybe_unused]] const double bbn = bn * bn; \
    [[maybe_unused]] const double DLM = SL - SM, DRM = SR - SM; \
    [[maybe_unused]] const double iLM = fast_rcp(DLM), iRM = fast_rcp(DRM); \
    [[maybe_unused]] const double ro_LL = ro_L * suL * iLM; \
    [[maybe_unused]] const double ro_RR = ro_R * suR * iRM;

// Reverse-mode hlld_face: the forward values recomputed with hlld_face's
// expressions (refined hardware reciprocal and square root, multiply-adds
// contracted within an expression, the contact speed's numerator products
// rounded on their own so mirror-image states give SM = 0 exactly), then one
// reverse sweep taking the cotangent Pb of the face-frame flux to the
// cotangents ib of the 16 face-frame inputs.
template <int METOD>
__device__ __forceinline__ void hlld_face_vjp(const double (&in)[16], double gamma, double inv_g1,
                                              const double (&Pb)[8], double (&ib)[16]) {
    // adjoints of every forward quantity
    double b_ro_L = 0.0, b_p_L = 0.0, b_u1 = 0.0, b_v1 = 0.0, b_w1 = 0.0, b_bn1 = 0.0, b_bt1 = 0.0, b_bm1 = 0.0;
    double b_ro_R = 0.0, b_p_R = 0.0, b_u2 = 0.0, b_v2 = 0.0, b_w2 = 0.0, b_bn2 = 0.0, b_bt2 = 0.0, b_bm2 = 0.0;
    // The cotangents of the two sides' fluxes F and states U are formed at
    // the end, from: the side the region formula takes (sel_left), the factor
    // of its state cotangent (sel_u: -S for HLLC-MHD; -S1, or 0 in the outer
    // region, for HLLD), the cotangents of the HLL field averages' F terms
    // (g_bt, g_bm) and of HLLC-MHD's HLL state UZ (gUZ): no 32-entry arrays
    // live through the sweep.
    bool sel_left = false;
    double sel_u = 0.0, g_bt = 0.0, g_bm = 0.0;
    double gUZ[4] = {0.0, 0.0, 0.0, 0.0};
    double b_SL = 0.0, b_SR = 0.0, b_SM = 0.0, b_bn = 0.0, b_bt = 0.0, b_bm = 0.0, b_ro_LL = 0.0, b_ro_RR = 0.0;
    double b_iS = 0.0, b_iLM = 0.0, b_iRM = 0.0, b_e1 = 0.0, b_e2 = 0.0, b_pTL = 0.0, b_pTR = 0.0;
    double b_suL = 0.0, b_suR = 0.0, b_mL = 0.0, b_mR = 0.0, b_idm = 0.0;

    {
    GD2_HLLD_VJP_FORWARD(in)
    // P[5] = -(smax * (bn2 - bn1)) replaces the region formula's P[5]
    {
        const double g = Pb[5];
        const double b_smax = -g * (bn2 - bn1);
        b_bn2 -= g * smax;
        b_bn1 += g * smax;
        const double b_a = smL ? b_smax : 0.0, b_b = smL ? 0.0 : b_smax;
        b_SL += SL < 0.0 ? -b_a : b_a;
        b_SR += SR < 0.0 ? -b_b : b_b;
    }

    if constexpr (METOD == 2) {
        double nUZ[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) nUZ[k] = SR * UR[k] - SL * UL[k] + FL[k] - FR[k];
        const double UZ0 = nUZ[0] * iS, UZ1 = nUZ[1] * iS, UZ2 = nUZ[2] * iS, UZ3 = nUZ[3] * iS;
        const double suRm = suR * iRM, suLm = suL * iLM;
        const double ptR = pTR + ro_R * suR * (SM - u2), ptL = pTL + ro_L * suL * (SM - u1);
        const double ptz = 0.5 * (ptR + ptL);
        const double iUZ0 = fast_rcp(UZ0);
        const double vv = UZ2 * iUZ0, ww = UZ3 * iUZ0;
        const double nz = bn * UZ1 + bt * UZ2 + bm * UZ3;
        const double sbvz = nz * iUZ0;
        const double tR = ptz * SM - pTR * u2 + bn * (sbv2 - sbvz);
        const double tL = ptz * SM - pTL * u1 + bn * (sbv1 - sbvz);
        const bool left = (SL >= 0.0) || (SM >= 0.0);
        const int sS = (SL >= 0.0) ? 0 : (left ? 1 : (SR > 0.0 ? 2 : 0));  // S = 0, SL, SR
        const double S = sS == 1 ? SL : (sS == 2 ? SR : 0.0);
        const double rs = left ? ro_LL : ro_RR;
        const double es = left ? e1 * suLm + tL * iLM : e2 * suRm + tR * iRM;
        const double Us[8] = {rs, rs * SM, rs * vv, rs * ww, es, bn, bt, bm};
        // P[i] = F[i] + S*Us[i] - S*U[i] (i != 5)
        double b_S = 0.0, bUs[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double g = i == 5 ? 0.0 : Pb[i];
            const double Ui = left ? UL[i] : UR[i];
            b_S += g * (Us[i] - Ui);
            bUs[i] = g * S;
        }
        sel_left = left;
        sel_u = -S;
        if (sS == 1) b_SL += b_S; else if (sS == 2) b_SR += b_S;
        double b_rs = bUs[0] + bUs[1] * SM + bUs[2] * vv + bUs[3] * ww;
        b_SM += bUs[1] * rs;
        const double b_vv = bUs[2] * rs, b_ww = bUs[3] * rs;
        const double b_es = bUs[4];
        b_bn += bUs[5];
        b_bt += bUs[6];
        b_bm += bUs[7];
        if (left) b_ro_LL += b_rs; else b_ro_RR += b_rs;
        double b_ptz = 0.0, b_sbvz = 0.0;
        if (left) {
            // es = e1*suLm + tL*iLM
            b_e1 += b_es * suLm;
            const double b_suLm = b_es * e1;
            const double b_tL = b_es * iLM;
            b_iLM += b_es * tL;
            b_suL += b_suLm * iLM;
            b_iLM += b_suLm * suL;
            // tL = ptz*SM - pTL*u1 + bn*(sbv1 - sbvz)
            b_ptz += b_tL * SM;
            b_SM += b_tL * ptz;
            b_pTL -= b_tL * u1;
            b_u1 -= b_tL * pTL;
            b_bn += b_tL * (sbv1 - sbvz);
            const double b_sb = b_tL * bn;
            b_sbvz -= b_sb;
            b_u1 += b_sb * bn1; b_bn1 += b_sb * u1;
            b_v1 += b_sb * bt1; b_bt1 += b_sb * v1;
            b_w1 += b_sb * bm1; b_bm1 += b_sb * w1;
        } else {
            b_e2 += b_es * suRm;
            const double b_suRm = b_es * e2;
            const double b_tR = b_es * iRM;
            b_iRM += b_es * tR;
            b_suR += b_suRm * iRM;
            b_iRM += b_suRm * suR;
            b_ptz += b_tR * SM;
            b_SM += b_tR * ptz;
            b_pTR -= b_tR * u2;
            b_u2 -= b_tR * pTR;
            b_bn += b_tR * (sbv2 - sbvz);
            const double b_sb = b_tR * bn;
            b_sbvz -= b_sb;
            b_u2 += b_sb * bn2; b_bn2 += b_sb * u2;
            b_v2 += b_sb * bt2; b_bt2 += b_sb * v2;
            b_w2 += b_sb * bm2; b_bm2 += b_sb * w2;
        }
        // sbvz = nz * iUZ0; vv = UZ2*iUZ0; ww = UZ3*iUZ0
        const double b_nz = b_sbvz * iUZ0;
        double b_iUZ0 = b_sbvz * nz + b_vv * UZ2 + b_ww * UZ3;
        double b_UZ2 = b_vv * iUZ0, b_UZ3 = b_ww * iUZ0;
        b_bn += b_nz * UZ1; b_bt += b_nz * UZ2; b_bm += b_nz * UZ3;
        const double b_UZ1 = b_nz * bn;
        b_UZ2 += b_nz * bt;
        b_UZ3 += b_nz * bm;
        const double b_UZ0 = -b_iUZ0 * iUZ0 * iUZ0;
        // ptz = 0.5*(ptR + ptL)
        const double b_ptR = 0.5 * b_ptz, b_ptL = 0.5 * b_ptz;
        b_pTR += b_ptR;
        {   // ro_R*suR*(SM-u2): ((ro_R*suR)*(SM-u2))
            const double a = ro_R * suR, d = SM - u2;
            const double b_a = b_ptR * d, b_d = b_ptR * a;
            b_ro_R += b_a * suR; b_suR += b_a * ro_R;
            b_SM += b_d; b_u2 -= b_d;
        }
        b_pTL += b_ptL;
        {
            const double a = ro_L * suL, d = SM - u1;
            const double b_a = b_ptL * d, b_d = b_ptL * a;
            b_ro_L += b_a * suL; b_suL += b_a * ro_L;
            b_SM += b_d; b_u1 -= b_d;
        }
        // UZk = (SR*UR[k] - SL*UL[k] + FL[k] - FR[k]) * iS
        const double bUZ[4] = {b_UZ0, b_UZ1, b_UZ2, b_UZ3};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double g = bUZ[k] * iS;
            b_iS += bUZ[k] * nUZ[k];
            b_SR += g * UR[k];
            b_SL -= g * UL[k];
            gUZ[k] = g;
        }
    } else {
        // ---------------------------------------------------------- HLLD forward
        const double a1 = ro_L * ro_R, a2 = a1 * suR, a3 = a2 * suL, du = u2 - u1;
        const double nP = mR * pTL - mL * pTR + a3 * du;
        const double PTT = nP * idm;
        const double kL = ro_L * suL, kR = ro_R * suR;
        const double ttL = kL * DLM - bbn, ttR = kR * DRM - bbn;
        const bool gL = fabs(ttL) >= 0.00001, gR = fabs(ttR) >= 0.00001;
        const double itL = gL ? fast_rcp(ttL) : 0.0, itR = gR ? fast_rcp(ttR) : 0.0;
        const double dSL = SM - u1, dSR = SM - u2;
        // vLL = v1 - ((bn*bt1)*dSL)*itL
        const double hvL = bn * bt1, hwL = bn * bm1, hvR = bn * bt2, hwR = bn * bm2;
        const double qnL = kL * suL - bbn, qnR = kR * suR - bbn;  // (ro*su)*su - bbn
        const double qL = qnL * itL, qR = qnR * itR;
        const double vLL = gL ? v1 - hvL * dSL * itL : v1, wLL = gL ? w1 - hwL * dSL * itL : w1;
        const double vRR = gR ? v2 - hvR * dSR * itR : v2, wRR = gR ? w2 - hwR * dSR * itR : w2;
        const double btt1 = gL ? bt1 * qL : 0.0, bmm1 = gL ? bm1 * qL : 0.0;
        const double btt2 = gR ? bt2 * qR : 0.0, bmm2 = gR ? bm2 * qR : 0.0;
        const double vbL = SM * bn + vLL * btt1 + wLL * bmm1;
        const double vbR = SM * bn + vRR * btt2 + wRR * bmm2;
        const double nEL = e1 * suL + PTT * SM - pTL * u1 + bn * (sbv1 - vbL);
        const double nER = e2 * suR + PTT * SM - pTR * u2 + bn * (sbv2 - vbR);
        const double eLL = nEL * iLM, eRR = nER * iRM;
        const double sLL = fast_sqrt(ro_LL), sRR = fast_sqrt(ro_RR);
        const double isLL = fast_rcp(sLL), isRR = fast_rcp(sRR);
        const double abn = bn < 0.0 ? -bn : bn;
        const bool big = fabs(bn) > 0.000001;
        const double idbn = big ? (bn > 0.0 ? 1.0 : -1.0) : 0.0;
        const double SLL = big ? SM - abn * isLL : SM;
        const double SRR = big ? SM + abn * isRR : SM;
        const double ssum = sLL + sRR;
        const double isum = fast_rcp(ssum);
        const double sLR = sLL * sRR;
        const double n1 = sLL * vLL + sRR * vRR + idbn * (btt2 - btt1);
        const double n2 = sLL * wLL + sRR * wRR + idbn * (bmm2 - bmm1);
        const double n3 = sLL * btt2 + sRR * btt1 + idbn * sLR * (vRR - vLL);
        const double n4 = sLL * bmm2 + sRR * bmm1 + idbn * sLR * (wRR - wLL);
        const double vLLL = n1 * isum, wLLL = n2 * isum, bttt = n3 * isum, bmmm = n4 * isum;
        const double vb3 = SM * bn + vLLL * bttt + wLLL * bmmm;
        const double eLLL = eLL - idbn * sLL * (vbL - vb3);
        const double eRRR = eRR + idbn * sRR * (vbR - vb3);
        const bool left = SM >= 0.0;
        const double S1 = left ? SL : SR, S2 = left ? SLL : SRR;
        const double rs = left ? ro_LL : ro_RR;
        const double vs = left ? vLL : vRR, ws = left ? wLL : wRR;
        const double bts = left ? btt1 : btt2, bms = left ? bmm1 : bmm2;
        const double es = left ? eLL : eRR, ess = left ? eLLL : eRRR;
        const bool outer = left ? (SL >= 0.0) : !(SRR > 0.0 || SR > 0.0);
        const bool single = left ? (SLL >= 0.0) : !(SRR > 0.0);
        const double Us[8] = {rs, rs * SM, rs * vs, rs * ws, es, bn, bts, bms};
        const double Uss[8] = {rs, rs * SM, rs * vLLL, rs * wLLL, ess, bn, bttt, bmmm};
        // ---------------------------------------------------------- HLLD reverse
        double bUs[8], bUss[8], b_S1 = 0.0, b_S2 = 0.0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const double g = i == 5 ? 0.0 : Pb[i];
            const double Ui = left ? UL[i] : UR[i];
            // region weights: outer F; single F + S1 (Us - U); double
            // F + S2 Uss - (S2 - S1) Us - S1 U
            const double w1 = outer ? 0.0 : 1.0, w2 = (outer || single) ? 0.0 : 1.0;
            b_S1 += g * w1 * (Us[i] - Ui);
            b_S2 += g * w2 * (Uss[i] - Us[i]);
            bUss[i] = g * S2 * w2;
            bUs[i] = w2 > 0.0 ? -g * (S2 - S1) : g * S1 * w1;
        }
        sel_left = left;
        sel_u = outer ? 0.0 : -S1;
        // Us = {rs, rs*SM, rs*vs, rs*ws, es, bn, bts, bms}
        double b_rs = bUs[0] + bUs[1] * SM + bUs[2] * vs + bUs[3] * ws;
        b_SM += bUs[1] * rs;
        const double b_vs = bUs[2] * rs, b_ws = bUs[3] * rs;
        const double b_es = bUs[4];
        b_bn += bUs[5];
        const double b_bts = bUs[6], b_bms = bUs[7];
        // Uss = {rs, rs*SM, rs*vLLL, rs*wLLL, ess, bn, bttt, bmmm}
        b_rs += bUss[0] + bUss[1] * SM + bUss[2] * vLLL + bUss[3] * wLLL;
        b_SM += bUss[1] * rs;
        double b_vLLL = bUss[2] * rs, b_wLLL = bUss[3] * rs;
        const double b_ess = bUss[4];
        b_bn += bUss[5];
        double b_bttt = bUss[6], b_bmmm = bUss[7];
        // route the side selections
        double b_vLL = 0.0, b_wLL = 0.0, b_vRR = 0.0, b_wRR = 0.0;
        double b_btt1 = 0.0, b_bmm1 = 0.0, b_btt2 = 0.0, b_bmm2 = 0.0;
        double b_eLL = 0.0, b_eRR = 0.0, b_eLLL = 0.0, b_eRRR = 0.0, b_SLL = 0.0, b_SRR = 0.0;
        if (left) {
            b_SL += b_S1; b_SLL += b_S2; b_ro_LL += b_rs;
            b_vLL += b_vs; b_wLL += b_ws; b_btt1 += b_bts; b_bmm1 += b_bms;
            b_eLL += b_es; b_eLLL += b_ess;
        } else {
            b_SR += b_S1; b_SRR += b_S2; b_ro_RR += b_rs;
            b_vRR += b_vs; b_wRR += b_ws; b_btt2 += b_bts; b_bmm2 += b_bms;
            b_eRR += b_es; b_eRRR += b_ess;
        }
        // eRRR = eRR + ((idbn*sRR)*(vbR - vb3)); eLLL = eLL - ((idbn*sLL)*(vbL - vb3))
        double b_vbL = 0.0, b_vbR = 0.0, b_vb3 = 0.0, b_sLL = 0.0, b_sRR = 0.0;
        b_eRR += b_eRRR;
        b_sRR += b_eRRR * idbn * (vbR - vb3);
        b_vbR += b_eRRR * idbn * sRR;
        b_vb3 -= b_eRRR * idbn * sRR;
        b_eLL += b_eLLL;
        b_sLL -= b_eLLL * idbn * (vbL - vb3);
        b_vbL -= b_eLLL * idbn * sLL;
        b_vb3 += b_eLLL * idbn * sLL;
        // vb3 = SM*bn + vLLL*bttt + wLLL*bmmm
        b_SM += b_vb3 * bn; b_bn += b_vb3 * SM;
        b_vLLL += b_vb3 * bttt; b_bttt += b_vb3 * vLLL;
        b_wLLL += b_vb3 * bmmm; b_bmmm += b_vb3 * wLLL;
        // vLLL = n1*isum ...
        const double b_n1 = b_vLLL * isum, b_n2 = b_wLLL * isum, b_n3 = b_bttt * isum, b_n4 = b_bmmm * isum;
        const double b_isum = b_vLLL * n1 + b_wLLL * n2 + b_bttt * n3 + b_bmmm * n4;
        const double b_ssum = -b_isum * isum * isum;
        b_sLL += b_ssum; b_sRR += b_ssum;
        // n1 = sLL*vLL + sRR*vRR + idbn*(btt2 - btt1)
        b_sLL += b_n1 * vLL; b_vLL += b_n1 * sLL;
        b_sRR += b_n1 * vRR; b_vRR += b_n1 * sRR;
        b_btt2 += b_n1 * idbn; b_btt1 -= b_n1 * idbn;
        b_sLL += b_n2 * wLL; b_wLL += b_n2 * sLL;
        b_sRR += b_n2 * wRR; b_wRR += b_n2 * sRR;
        b_bmm2 += b_n2 * idbn; b_bmm1 -= b_n2 * idbn;
        // n3 = sLL*btt2 + sRR*btt1 + ((idbn*sLR)*(vRR - vLL))
        double b_sLR = 0.0;
        b_sLL += b_n3 * btt2; b_btt2 += b_n3 * sLL;
        b_sRR += b_n3 * btt1; b_btt1 += b_n3 * sRR;
        b_sLR += b_n3 * idbn * (vRR - vLL);
        b_vRR += b_n3 * idbn * sLR; b_vLL -= b_n3 * idbn * sLR;
        b_sLL += b_n4 * bmm2; b_bmm2 += b_n4 * sLL;
        b_sRR += b_n4 * bmm1; b_bmm1 += b_n4 * sRR;
        b_sLR += b_n4 * idbn * (wRR - wLL);
        b_wRR += b_n4 * idbn * sLR; b_wLL -= b_n4 * idbn * sLR;
        b_sLL += b_sLR * sRR; b_sRR += b_sLR * sLL;
        // SLL = SM - abn*isLL, SRR = SM + abn*isRR (when big; else SM)
        b_SM += b_SLL + b_SRR;
        if (big) {
            const double b_abn = -b_SLL * isLL + b_SRR * isRR;
            const double b_isLL = -b_SLL * abn, b_isRR = b_SRR * abn;
            b_sLL -= b_isLL * isLL * isLL;
            b_sRR -= b_isRR * isRR * isRR;
            b_bn += bn < 0.0 ? -b_abn : b_abn;
        }
        // sLL = sqrt(ro_LL)
        b_ro_LL += sLL > 0.0 ? 0.5 * b_sLL * isLL : 0.0;
        b_ro_RR += sRR > 0.0 ? 0.5 * b_sRR * isRR : 0.0;
        // eLL = nEL*iLM, nEL = e1*suL + PTT*SM - pTL*u1 + bn*(sbv1 - vbL)
        double b_PTT = 0.0;
        {
            const double b_n = b_eLL * iLM;
            b_iLM += b_eLL * nEL;
            b_e1 += b_n * suL; b_suL += b_n * e1;
            b_PTT += b_n * SM; b_SM += b_n * PTT;
            b_pTL -= b_n * u1; b_u1 -= b_n * pTL;
            b_bn += b_n * (sbv1 - vbL);
            const double b_sb = b_n * bn;
            b_vbL -= b_sb;
            b_u1 += b_sb * bn1; b_bn1 += b_sb * u1;
            b_v1 += b_sb * bt1; b_bt1 += b_sb * v1;
            b_w1 += b_sb * bm1; b_bm1 += b_sb * w1;
        }
        {
            const double b_n = b_eRR * iRM;
            b_iRM += b_eRR * nER;
            b_e2 += b_n * suR; b_suR += b_n * e2;
            b_PTT += b_n * SM; b_SM += b_n * PTT;
            b_pTR -= b_n * u2; b_u2 -= b_n * pTR;
            b_bn += b_n * (sbv2 - vbR);
            const double b_sb = b_n * bn;
            b_vbR -= b_sb;
            b_u2 += b_sb * bn2; b_bn2 += b_sb * u2;
            b_v2 += b_sb * bt2; b_bt2 += b_sb * v2;
            b_w2 += b_sb * bm2; b_bm2 += b_sb * w2;
        }
        // vbL = SM*bn + vLL*btt1 + wLL*bmm1
        b_SM += b_vbL * bn + b_vbR * bn;
        b_bn += b_vbL * SM + b_vbR * SM;
        b_vLL += b_vbL * btt1; b_btt1 += b_vbL * vLL;
        b_wLL += b_vbL * bmm1; b_bmm1 += b_vbL * wLL;
        b_vRR += b_vbR * btt2; b_btt2 += b_vbR * vRR;
        b_wRR += b_vbR * bmm2; b_bmm2 += b_vbR * wRR;
        // the tangential star states
        double b_kL = 0.0, b_kR = 0.0, b_bbn = 0.0;
        if (gL) {
            // btt1 = bt1*qL, bmm1 = bm1*qL; qL = qnL*itL; vLL = v1 - ((hvL*dSL)*itL); wLL likewise
            b_bt1 += b_btt1 * qL; b_bm1 += b_bmm1 * qL;
            const double b_qL = b_btt1 * bt1 + b_bmm1 * bm1;
            const double b_qn = b_qL * itL;
            double b_it = b_qL * qnL;
            b_kL += b_qn * suL; b_suL += b_qn * kL; b_bbn -= b_qn;
            b_v1 += b_vLL; b_w1 += b_wLL;
            const double gv = -b_vLL, gw = -b_wLL;
            b_it += gv * hvL * dSL + gw * hwL * dSL;
            const double b_dS = gv * hvL * itL + gw * hwL * itL;
            const double b_hv = gv * dSL * itL, b_hw = gw * dSL * itL;
            b_bn += b_hv * bt1 + b_hw * bm1;
            b_bt1 += b_hv * bn; b_bm1 += b_hw * bn;
            b_SM += b_dS; b_u1 -= b_dS;
            const double b_tt = -b_it * itL * itL;
            b_kL += b_tt * DLM;
            const double b_DLM = b_tt * kL;
            b_SL += b_DLM; b_SM -= b_DLM;
            b_bbn -= b_tt;
        } else {
            b_v1 += b_vLL; b_w1 += b_wLL;
        }
        if (gR) {
            b_bt2 += b_btt2 * qR; b_bm2 += b_bmm2 * qR;
            const double b_qR = b_btt2 * bt2 + b_bmm2 * bm2;
            const double b_qn = b_qR * itR;
            double b_it = b_qR * qnR;
            b_kR += b_qn * suR; b_suR += b_qn * kR; b_bbn -= b_qn;
            b_v2 += b_vRR; b_w2 += b_wRR;
            const double gv = -b_vRR, gw = -b_wRR;
            b_it += gv * hvR * dSR + gw * hwR * dSR;
            const double b_dS = gv * hvR * itR + gw * hwR * itR;
            const double b_hv = gv * dSR * itR, b_hw = gw * dSR * itR;
            b_bn += b_hv * bt2 + b_hw * bm2;
            b_bt2 += b_hv * bn; b_bm2 += b_hw * bn;
            b_SM += b_dS; b_u2 -= b_dS;
            const double b_tt = -b_it * itR * itR;
            b_kR += b_tt * DRM;
            const double b_DRM = b_tt * kR;
            b_SR += b_DRM; b_SM -= b_DRM;
            b_bbn -= b_tt;
        } else {
            b_v2 += b_vRR; b_w2 += b_wRR;
        }
        b_ro_L += b_kL * suL; b_suL += b_kL * ro_L;
        b_ro_R += b_kR * suR; b_suR += b_kR * ro_R;
        b_bn += 2.0 * bn * b_bbn;
        // PTT = nP*idm, nP = mR*pTL - mL*pTR + ((ro_L*ro_R)*suR*suL)*(u2 - u1)
        {
            const double b_n = b_PTT * idm;
            b_idm += b_PTT * nP;
            b_mR += b_n * pTL; b_pTL += b_n * mR;
            b_mL -= b_n * pTR; b_pTR -= b_n * mL;
            const double b_a3 = b_n * du, b_du = b_n * a3;
            b_u2 += b_du; b_u1 -= b_du;
            const double b_a2 = b_a3 * suL; b_suL += b_a3 * a2;
            const double b_a1 = b_a2 * suR; b_suR += b_a2 * a1;
            b_ro_L += b_a1 * ro_R; b_ro_R += b_a1 * ro_L;
        }
    }
    }
    // The common part of the sweep reads the forward quantities again,
    // recomputed from the inputs passed through an empty asm: the compiler
    // cannot reuse the first copies, so those need not stay live through the
    // flux-specific part (the recomputation is bitwise the first).
    double in2[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        double x = in[i];
        asm volatile("" : "+v"(x));
        in2[i] = x;
    }
    GD2_HLLD_VJP_FORWARD(in2)
    // ---------------------------------------------------------------- common reverse
    // bbn = bn*bn is folded above for HLLD; HLLC-MHD does not use bbn
    // ro_LL = (ro_L*suL)*iLM, ro_RR likewise
    {
        const double k = ro_L * suL;
        b_iLM += b_ro_LL * k;
        const double b_k = b_ro_LL * iLM;
        b_ro_L += b_k * suL; b_suL += b_k * ro_L;
    }
    {
        const double k = ro_R * suR;
        b_iRM += b_ro_RR * k;
        const double b_k = b_ro_RR * iRM;
        b_ro_R += b_k * suR; b_suR += b_k * ro_R;
    }
    // iLM = 1/(SL - SM), iRM = 1/(SR - SM)
    {
        const double g = -b_iLM * iLM * iLM;
        b_SL += g; b_SM -= g;
        const double h = -b_iRM * iRM * iRM;
        b_SR += h; b_SM -= h;
    }
    // bn = nbn*iS, bt = nbt*iS, bm = nbm*iS
    {
        const double gn = b_bn * iS, gt = b_bt * iS, gm = b_bm * iS;
        b_iS += b_bn * nbn + b_bt * nbt + b_bm * nbm;
        b_SR += gn * bn2 + gt * bt2 + gm * bm2;
        b_SL -= gn * bn1 + gt * bt1 + gm * bm1;
        b_bn2 += gn * SR; b_bn1 -= gn * SL;
        b_bt2 += gt * SR; b_bt1 -= gt * SL;
        b_bm2 += gm * SR; b_bm1 -= gm * SL;
        g_bt = gt;
        g_bm = gm;
        const double g = -b_iS * iS * iS;
        b_SR += g; b_SL -= g;
    }
    // U and F of each side
    double bFL[8], bFR[8], bUL[8], bUR[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const double g = i == 5 ? 0.0 : Pb[i];
        const double gu = g * sel_u;
        bFL[i] = sel_left ? g : 0.0;
        bFR[i] = sel_left ? 0.0 : g;
        bUL[i] = sel_left ? gu : 0.0;
        bUR[i] = sel_left ? 0.0 : gu;
    }
    if constexpr (METOD == 2) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            bUR[k] += gUZ[k] * SR;
            bUL[k] -= gUZ[k] * SL;
            bFL[k] += gUZ[k];
            bFR[k] -= gUZ[k];
        }
    }
    bFL[6] += g_bt;
    bFR[6] -= g_bt;
    bFL[7] += g_bm;
    bFR[7] -= g_bm;
    b_ro_L += bUL[0] + bUL[1] * u1 + bUL[2] * v1 + bUL[3] * w1;
    b_u1 += bUL[1] * ro_L; b_v1 += bUL[2] * ro_L; b_w1 += bUL[3] * ro_L;
    b_e1 += bUL[4]; b_bn1 += bUL[5]; b_bt1 += bUL[6]; b_bm1 += bUL[7];
    b_ro_R += bUR[0] + bUR[1] * u2 + bUR[2] * v2 + bUR[3] * w2;
    b_u2 += bUR[1] * ro_R; b_v2 += bUR[2] * ro_R; b_w2 += bUR[3] * ro_R;
    b_e2 += bUR[4]; b_bn2 += bUR[5]; b_bt2 += bUR[6]; b_bm2 += bUR[7];
    // FL: [0] ro u, [1] (ro u) u + pT - bn^2, [2] (ro u) v - bn bt, [3] (ro u) w - bn bm,
    // [4] (e + pT) u - bn sbv, [6] u bt - v bn, [7] u bm - w bn
    {
        const double ru = ro_L * u1;
        const double b_ru = bFL[0] + bFL[1] * u1 + bFL[2] * v1 + bFL[3] * w1;
        b_ro_L += b_ru * u1; b_u1 += b_ru * ro_L;
        b_u1 += bFL[1] * ru + bFL[4] * (e1 + pTL) + bFL[6] * bt1 + bFL[7] * bm1;
        b_pTL += bFL[1] + bFL[4] * u1;
        b_bn1 -= 2.0 * bn1 * bFL[1] + bFL[2] * bt1 + bFL[3] * bm1 + bFL[4] * sbv1 + bFL[6] * v1 + bFL[7] * w1;
        b_v1 += bFL[2] * ru - bFL[6] * bn1;
        b_w1 += bFL[3] * ru - bFL[7] * bn1;
        b_bt1 += -bFL[2] * bn1 + bFL[6] * u1;
        b_bm1 += -bFL[3] * bn1 + bFL[7] * u1;
        b_e1 += bFL[4] * u1;
        const double b_sb = -bFL[4] * bn1;
        b_u1 += b_sb * bn1; b_bn1 += b_sb * u1;
        b_v1 += b_sb * bt1; b_bt1 += b_sb * v1;
        b_w1 += b_sb * bm1; b_bm1 += b_sb * w1;
    }
    {
        const double ru = ro_R * u2;
        const double b_ru = bFR[0] + bFR[1] * u2 + bFR[2] * v2 + bFR[3] * w2;
        b_ro_R += b_ru * u2; b_u2 += b_ru * ro_R;
        b_u2 += bFR[1] * ru + bFR[4] * (e2 + pTR) + bFR[6] * bt2 + bFR[7] * bm2;
        b_pTR += bFR[1] + bFR[4] * u2;
        b_bn2 -= 2.0 * bn2 * bFR[1] + bFR[2] * bt2 + bFR[3] * bm2 + bFR[4] * sbv2 + bFR[6] * v2 + bFR[7] * w2;
        b_v2 += bFR[2] * ru - bFR[6] * bn2;
        b_w2 += bFR[3] * ru - bFR[7] * bn2;
        b_bt2 += -bFR[2] * bn2 + bFR[6] * u2;
        b_bm2 += -bFR[3] * bn2 + bFR[7] * u2;
        b_e2 += bFR[4] * u2;
        const double b_sb = -bFR[4] * bn2;
        b_u2 += b_sb * bn2; b_bn2 += b_sb * u2;
        b_v2 += b_sb * bt2; b_bt2 += b_sb * v2;
        b_w2 += b_sb * bm2; b_bm2 += b_sb * w2;
    }
    // e = p*inv_g1 + ro*uu + 0.5*bb
    double b_bb_L = 0.5 * b_e1, b_bb_R = 0.5 * b_e2, b_uu_L = b_e1 * ro_L, b_uu_R = b_e2 * ro_R;
    b_p_L += b_e1 * inv_g1; b_ro_L += b_e1 * uu_L;
    b_p_R += b_e2 * inv_g1; b_ro_R += b_e2 * uu_R;
    // SM = Nn*idm, Nn = mR*u2 - mL*u1 - pTR + pTL
    {
        const double b_N = b_SM * idm;
        b_idm += b_SM * Nn;
        b_mR += b_N * u2; b_u2 += b_N * mR;
        b_mL -= b_N * u1; b_u1 -= b_N * mL;
        b_pTR -= b_N; b_pTL += b_N;
        const double b_D = -b_idm * idm * idm;
        b_mR += b_D; b_mL -= b_D;
    }
    // mR = suR*ro_R, mL = suL*ro_L; suR = SR - u2, suL = SL - u1
    b_suR += b_mR * ro_R; b_ro_R += b_mR * suR;
    b_suL += b_mL * ro_L; b_ro_L += b_mL * suL;
    b_SR += b_suR; b_u2 -= b_suR;
    b_SL += b_suL; b_u1 -= b_suL;
    // pT = p + 0.5*bb
    b_p_L += b_pTL; b_bb_L += 0.5 * b_pTL;
    b_p_R += b_pTR; b_bb_R += 0.5 * b_pTR;
    // SL = umin - cmax, SR = umax + cmax
    const double b_cmax = b_SR - b_SL;
    if (uminL) b_u1 += b_SL; else b_u2 += b_SL;
    if (umaxL) b_u1 += b_SR; else b_u2 += b_SR;
    double b_cfL = cL ? b_cmax : 0.0, b_cfR = cL ? 0.0 : b_cmax;
    // cf = sqrt(a), a = (gp + sqrt(d)) * (0.5*iro), d = gp^2 - (4 gamma p) bn^2
    {
        const double b_a = cfL > 0.0 ? 0.5 * b_cfL * fast_rcp(cfL) : 0.0;
        const double h = 0.5 * iroL;
        const double b_gs = b_a * h;
        const double b_iro = b_a * (gpL + sdL) * 0.5;
        const double b_d = sdL > 0.0 ? 0.5 * b_gs * fast_rcp(sdL) : 0.0;
        const double b_gp = b_gs + b_d * 2.0 * gpL;
        // d = gp*gp - (g4*p)*(bn^2)
        b_p_L -= b_d * g4 * (bn1 * bn1);
        b_bn1 -= b_d * (g4 * p_L) * 2.0 * bn1;
        b_p_L += b_gp * gamma;
        b_bb_L += b_gp;
        b_ro_L -= b_iro * iroL * iroL;
    }
    {
        const double b_a = cfR > 0.0 ? 0.5 * b_cfR * fast_rcp(cfR) : 0.0;
        const double h = 0.5 * iroR;
        const double b_gs = b_a * h;
        const double b_iro = b_a * (gpR + sdR) * 0.5;
        const double b_d = sdR > 0.0 ? 0.5 * b_gs * fast_rcp(sdR) : 0.0;
        const double b_gp = b_gs + b_d * 2.0 * gpR;
        b_p_R -= b_d * g4 * (bn2 * bn2);
        b_bn2 -= b_d * (g4 * p_R) * 2.0 * bn2;
        b_p_R += b_gp * gamma;
        b_bb_R += b_gp;
        b_ro_R -= b_iro * iroR * iroR;
    }
    // uu = 0.5*(u^2 + v^2 + w^2), bb = bn^2 + bt^2 + bm^2
    b_u1 += b_uu_L * u1; b_v1 += b_uu_L * v1; b_w1 += b_uu_L * w1;
    b_u2 += b_uu_R * u2; b_v2 += b_uu_R * v2; b_w2 += b_uu_R * w2;
    b_bn1 += 2.0 * b_bb_L * bn1; b_bt1 += 2.0 * b_bb_L * bt1; b_bm1 += 2.0 * b_bb_L * bm1;
    b_bn2 += 2.0 * b_bb_R * bn2; b_bt2 += 2.0 * b_bb_R * bt2; b_bm2 += 2.0 * b_bb_R * bm2;
    ib[0] = b_ro_L; ib[1] = b_p_L; ib[2] = b_u1; ib[3] = b_v1; ib[4] = b_w1; ib[5] = b_bn1; ib[6] = b_bt1; ib[7] = b_bm1;
    ib[8] = b_ro_R; ib[9] = b_p_R; ib[10] = b_u2; ib[11] = b_v2; ib[12] = b_w2; ib[13] = b_bn2; ib[14] = b_bt2;
    ib[15] = b_bm2;
}

// mhd_flux reversed: L, R as r, u, v, p, bx, by, bz (Gaussian B); phi the
// cotangent of the lab-frame fluxes in conserved order r, mx, my, bx, by,
// bz, e; Lb, Rb the cotangents of L and R.
template <int METOD, bool XFACE>
__device__ __forceinline__ void mhd_flux_vjp(const double (&L)[7], const double (&R)[7], double gamma,
                                             double inv_g1, const double (&phi)[7], double (&Lb)[7],
                                             double (&Rb)[7]) {
    const double is = 1.0 / kSpi4;
    double in[16], Pb[8];
    if constexpr (XFACE) {
        const double a[16] = {L[0], L[3], L[1], L[2], 0.0, L[4] * is, L[5] * is, L[6] * is,
                              R[0], R[3], R[1], R[2], 0.0, R[4] * is, R[5] * is, R[6] * is};
#pragma unroll
        for (int i = 0; i < 16; ++i) in[i] = a[i];
        const double b[8] = {phi[0], phi[1], phi[2], 0.0, phi[6], kSpi4 * phi[3], kSpi4 * phi[4], kSpi4 * phi[5]};
#pragma unroll
        for (int i = 0; i < 8; ++i) Pb[i] = b[i];
    } else {
        const double a[16] = {L[0], L[3], L[2], 0.0, L[1], L[5] * is, L[6] * is, L[4] * is,
                              R[0], R[3], R[2], 0.0, R[1], R[5] * is, R[6] * is, R[4] * is};
#pragma unroll
        for (int i = 0; i < 16; ++i) in[i] = a[i];
        const double b[8] = {phi[0], phi[2], 0.0, phi[1], phi[6], kSpi4 * phi[4], kSpi4 * phi[5], kSpi4 * phi[3]};
#pragma unroll
        for (int i = 0; i < 8; ++i) Pb[i] = b[i];
    }
    double ib[16];
    hlld_face_vjp<METOD>(in, gamma, inv_g1, Pb, ib);
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
        const double* g = ib + 8 * sd;
        double* o = sd == 0 ? Lb : Rb;
        o[0] = g[0];
        o[3] = g[1];
        if constexpr (XFACE) {
            o[1] = g[2];
            o[2] = g[3];
            o[4] = g[5] * is;
            o[5] = g[6] * is;
            o[6] = g[7] * is;
        } else {
            o[2] = g[2];
            o[1] = g[4];
            o[5] = g[5] * is;
            o[6] = g[6] * is;
            o[4] = g[7] * is;
        }
    }
}

struct MhdAdjArgs {
    const double* U;  // state the stage operator is linearised at (7 SoA planes)
    const double* V;  // cotangent
    const double* A;  // optional
    const double* B;  // optional
    double* W;        // output
    double a, b, c;   // W = a*A + b*B + c*J^T V
    double* cx;       // x-gathers of the cells: 7 planes of nx x ny
    double* wy;       // y-face state cotangents: 14 planes of nx x (ny+1)
    double* gx;       // gathers of the ghost columns -2, -1, nx, nx+1: [4][ny][7]
    double* gy;       // gathers of the ghost rows -2, -1, ny, ny+1: [4][nx][7]
    int tiles_x, rows;  // strips of the gather march
    int nx, ny;
    long long plane;
    long long nfy;    // y-faces: nx x (ny+1)
    double dx, dy, dV, idV;  // idV = 1 / dV
    double hfx, hfy;         // dy / dV, dx / dV: x- and y-face length over cell area
    double gamma, g1, inv_g1;
    int bc[4];
};

// Primitives of cell (i, j), either index possibly outside the grid: the
// boundary map's source cell with the mirrored components negated, as the
// forward stage reads it.
__device__ __forceinline__ PrimM adj_cell_m(const MhdAdjArgs& J, int i, int j) {
    bool fx, fy;
    const int ii = map_index(i, J.nx, J.bc[0], J.bc[1], fx);
    const int jj = map_index(j, J.ny, J.bc[2], J.bc[3], fy);
    PrimM q = cons_to_prim_m(J.U, J.plane, (long long)jj * J.nx + ii, J.g1);
    if (fx) {
        q.u = -q.u;
        q.bx = -q.bx;
    }
    if (fy) {
        q.v = -q.v;
        q.by = -q.by;
    }
    return q;
}

__device__ __forceinline__ double comp_m(const PrimM& q, int f) {
    switch (f) {
        case 0: return q.r;
        case 1: return q.u;
        case 2: return q.v;
        case 3: return q.p;
        case 4: return q.bx;
        case 5: return q.by;
        default: return q.bz;
    }
}

// Primitives of a cell from its 7 conserved values (the expressions of
// cons_to_prim_m); the face march and the gather both convert through it.
__device__ __forceinline__ PrimM prim_of_cons7(const double (&w)[7], double g1) {
    PrimM q;
    const double r = w[0], mx = w[1], my = w[2];
    q.bx = w[3];
    q.by = w[4];
    q.bz = w[5];
    const double E = w[6];
    const double inv = fast_rcp(r);
    q.r = r;
    q.u = mx * inv;
    q.v = my * inv;
    q.p = g1 * (E - 0.5 * (mx * q.u + my * q.v) - (q.bx * q.bx + q.by * q.by + q.bz * q.bz) * (1.0 / kCpi8));
    return q;
}

// One cell's shares of the reconstruction adjoint along one line: qa, qb, qd
// the cell before, the cell itself and the cell after; dqp the cotangent of
// its plus-face value (left state of the face after it), dqm of its
// minus-face value (right state of the face before it).
template <int ORDER>
__device__ __forceinline__ void spread7(const PrimM& qa, const PrimM& qb, const PrimM& qd, const double (&dqp)[7],
                                        const double (&dqm)[7], double (&own)[7], double (&to_prev)[7],
                                        double (&to_next)[7]) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        if constexpr (ORDER == 1) {
            own[k] = dqp[k] + dqm[k];
            to_prev[k] = to_next[k] = 0.0;
        } else {
            const double a = comp_m(qa, k), b = comp_m(qb, k), d = comp_m(qd, k);
            const MMW wt = mm_weights(b - a, d - b);
            bool fp = false, fm = false;
            if (k == 0 || k == 3) {
                const double sl = minmod(b - a, d - b);
                fp = !(b + 0.5 * sl > 0.0);
                fm = !(b + -0.5 * sl > 0.0);
            }
            spread(dqp[k], dqm[k], wt, fp, fm, own[k], to_prev[k], to_next[k]);
        }
    }
}

// The 14 face-state cotangents of one face (planes 0..6: its left state,
// 7..13: its right state) from its reconstructed states L, R and its face
// cotangent phi (conserved order). A reflective west / south wall face is
// evaluated mirrored, F = unmirror(flux(mirror(R), mirror(L))), as the forward
// stage does.
template <int METOD, bool XFACE>
__device__ __forceinline__ void face_state_cot(const MhdAdjArgs& J, const PrimM& L, const PrimM& R, double (&phi)[7],
                                               bool wall, double (&s)[14]) {
    PrimM X = L, Y = R;
    if (wall) {
        X = XFACE ? mirror_x(R) : mirror_y(R);
        Y = XFACE ? mirror_x(L) : mirror_y(L);
        // unmirror negates mass, energy and the tangential momentum / field
        phi[0] = -phi[0];
        phi[6] = -phi[6];
        phi[5] = -phi[5];
        if (XFACE) {
            phi[2] = -phi[2];
            phi[4] = -phi[4];
        } else {
            phi[1] = -phi[1];
            phi[3] = -phi[3];
        }
    }
    const int mu = XFACE ? 1 : 2, mb = XFACE ? 4 : 5;  // components a mirror negates
    double Xv[7], Yv[7], Xb[7], Yb[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        Xv[c] = comp_m(X, c);
        Yv[c] = comp_m(Y, c);
    }
    mhd_flux_vjp<METOD, XFACE>(Xv, Yv, J.gamma, J.inv_g1, phi, Xb, Yb);
    // on a wall face X is the mirrored right state and Y the mirrored left
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const bool neg = c == mu || c == mb;
        s[c] = wall ? (neg ? -Yb[c] : Yb[c]) : Xb[c];
        s[c + 7] = wall ? (neg ? -Xb[c] : Xb[c]) : Yb[c];
    }
}

constexpr int kFaceW = 58;  // strip width of the face march (lanes 3..60 own its columns)

// Rows of a strip of the face march in LDS: the primitive rows p-1 and p
// (slot row & 1; row p+1 replaces row p-1 once the y-face has read it), the
// cotangent row p-1 and the upper-face value of row p-1 (the left state of
// its y-face).
struct MhdFaceLds {
    double q[2][kMF][kRowBuf];
    double v[1][kMF][kRowBuf];
    double up[1][kMF][kRowBuf];
};

// The face march: one wave per strip of kFaceW columns and rows [j0, j1).
// Lane l holds column c = cbase - 3 + l (cbase = kFaceW tx); lanes 3..60 own
// the strip's columns. Pass p loads row p+1 (through the boundary map) and
// reverses the flux of
//  * the y-face between rows p-1 and p (owned columns): its 14 face-state
//    cotangents go to the y-face buffer, for the gather march;
//  * the x-face c + 1/2 of row p (lanes 1..61; a face outside -1/2 .. nx-1/2
//    has none): the cotangents of its two states are spread over their
//    minmod stencils along the row at once, and the x-gather of each owned
//    cell (its own share and its neighbours' shares, by lane shifts) goes to
//    the cell's x-gather planes; the last strip writes the gathers of the
//    ghost columns nx, nx+1 to their side buffer, the first those of -2, -1.
// Passes run from j0 to j1 - 1, and to ny on the top strip (its y-face
// ny - 1/2). Each cell is loaded and converted once per strip; the face
// states come from the LDS rows and lane shifts, and nothing stays in
// registers across a face reversal (the empty asm statements keep the
// compiler from forwarding LDS values across one).
template <int ORDER, int METOD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void mhd_vjp_face_march_kernel(
    const MhdAdjArgs J) {
    __shared__ MhdFaceLds lds;
    const int lane = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % J.tiles_x, ty = tile / J.tiles_x;
    const int nx = J.nx, ny = J.ny;
    const int c = tx * kFaceW - 3 + lane;
    const int j0 = ty * J.rows, j1 = min(j0 + J.rows, ny);
    const int pe = j1 == ny ? ny : j1 - 1;  // last pass
    const bool col_in = c >= 0 && c < nx;
    const bool own = col_in && lane >= 3 && lane <= 60;
    const bool face_x = c >= -1 && c <= nx - 1;  // face c + 1/2 exists
    const bool wall_x = c == -1 && J.bc[0] == GD2_BC_REFLECTIVE;
    // ghost column of this lane whose x-gather this wave writes (0..3 for
    // -2, -1, nx, nx+1; -1 otherwise)
    int gslot = -1;
    if (c == -2 || c == -1) gslot = c + 2;
    else if ((c == nx || c == nx + 1) && tx == J.tiles_x - 1) gslot = c - nx + 2;
    bool fx;
    const int ci = map_index(c, nx, J.bc[0], J.bc[1], fx);
    const unsigned oc = (unsigned)ci * 8u;
    const unsigned ov = col_in ? (unsigned)c * 8u : kOffNone;  // V of this column
    const unsigned oo = own ? (unsigned)c * 8u : kOffNone;     // this cell's y-face and x-gather
    const long long P = J.plane;

    auto load_raw = [&](int jl, double (&w)[7]) {
        bool fy;
        const int jj = map_index(jl, ny, J.bc[2], J.bc[3], fy);
        const double* b = J.U + (long long)jj * nx;
#pragma unroll
        for (int f = 0; f < 7; ++f) w[f] = ld_buf(buf_rsrc(uniform_ptr(b + f * P)), oc);
        return fy;
    };
    auto to_prim = [&](const double (&w)[7], bool fy) -> PrimM {
        PrimM q = prim_of_cons7(w, J.g1);
        if (fx) {
            q.u = -q.u;
            q.bx = -q.bx;
        }
        if (fy) {
            q.v = -q.v;
            q.by = -q.by;
        }
        return q;
    };
    auto load_v = [&](int r, double (&v)[7]) {
        const bool in = r >= 0 && r < ny;
        const double* b = J.V + (long long)(in ? r : 0) * nx;
        const unsigned o = in ? ov : kOffNone;
#pragma unroll
        for (int f = 0; f < 7; ++f) v[f] = ld_buf(buf_rsrc(uniform_ptr(b + f * P)), o);
    };
    // the row's primitives at this lane and its two neighbours
    auto row_window = [&](int p, PrimM& qw, PrimM& q, PrimM& qe) {
        // the lane index taken afresh: LDS addresses are formed here, not
        // held across the face reversals
        int e = lane;
        asm volatile("" : "+v"(e));
        q = lds_get_m(lds.q, p & 1, e);
        qw.r = from_prev_lane(q.r); qe.r = from_next_lane(q.r);
        qw.u = from_prev_lane(q.u); qe.u = from_next_lane(q.u);
        qw.v = from_prev_lane(q.v); qe.v = from_next_lane(q.v);
        qw.p = from_prev_lane(q.p); qe.p = from_next_lane(q.p);
        qw.bx = from_prev_lane(q.bx); qe.bx = from_next_lane(q.bx);
        qw.by = from_prev_lane(q.by); qe.by = from_next_lane(q.by);
        qw.bz = from_prev_lane(q.bz); qe.bz = from_next_lane(q.bz);
    };

    // prologue: rows j0-1 and j0, the upper-face value of row j0-1 and the
    // cotangent row j0-1
    {
        double wa[7], wb[7], wc[7], v[7];
        const bool fa = load_raw(j0 - 2, wa), fb = load_raw(j0 - 1, wb), fc = load_raw(j0, wc);
        load_v(j0 - 1, v);
        const PrimM qa = to_prim(wa, fa), qb = to_prim(wb, fb), qc = to_prim(wc, fc);
        lds_put_m(lds.q, (j0 - 1) & 1, lane, qb);
        lds_put_m(lds.q, j0 & 1, lane, qc);
        lds_put_m(lds.up, 0, lane, face_m(qb, slope3m<ORDER>(qa, qb, qc), 0.5));
#pragma unroll
        for (int f = 0; f < 7; ++f) lds.v[0][f][lane] = v[f];
    }

    for (int p = j0; p <= pe; ++p) {
        asm volatile("" ::: "memory");
        {
            // y-face between rows p-1 and p; row p+1 then takes row p-1's slot
            // and the cotangent row p the slot of row p-1
            double w[7], vp[7];
            const bool fn = load_raw(p + 1, w);
            load_v(p, vp);
            int e = lane;
            asm volatile("" : "+v"(e));
            const PrimM qs = lds_get_m(lds.q, (p - 1) & 1, e), q = lds_get_m(lds.q, p & 1, e);
            const PrimM L = lds_get_m(lds.up, 0, e);
            double phi[7];
#pragma unroll
            for (int f = 0; f < 7; ++f) {
                const double vs = lds.v[0][f][e];
                phi[f] = J.hfy * (vp[f] - vs);
                lds.v[0][f][e] = vp[f];
            }
            const PrimM qn = to_prim(w, fn);
            lds_put_m(lds.q, (p + 1) & 1, e, qn);
            const PrimM s = slope3m<ORDER>(qs, q, qn);
            const PrimM R = face_m(q, s, -0.5);
            lds_put_m(lds.up, 0, e, face_m(q, s, 0.5));
            const bool wall = p == 0 && J.bc[2] == GD2_BC_REFLECTIVE;
            double o[14];
            face_state_cot<METOD, false>(J, L, R, phi, wall, o);
            double* wy = J.wy + (long long)p * nx;
#pragma unroll
            for (int k = 0; k < 14; ++k) st_buf(buf_rsrc(uniform_ptr(wy + k * J.nfy)), oo, o[k]);
        }
        asm volatile("" ::: "memory");
        if (p < j1) {
            // x-face c + 1/2 of row p: states from the row's lane shifts,
            // cotangent from this column's and the next column's V
            double xdqp[7], xdqm[7];
            {
                PrimM qw, q, qe;
                row_window(p, qw, q, qe);
                const PrimM s = slope3m<ORDER>(qw, q, qe);
                const PrimM L = face_m(q, s, 0.5), Rm = face_m(q, s, -0.5);
                PrimM R;
                R.r = from_next_lane(Rm.r);
                R.u = from_next_lane(Rm.u);
                R.v = from_next_lane(Rm.v);
                R.p = from_next_lane(Rm.p);
                R.bx = from_next_lane(Rm.bx);
                R.by = from_next_lane(Rm.by);
                R.bz = from_next_lane(Rm.bz);
                int e = lane;
                asm volatile("" : "+v"(e));
                double phi[7];
#pragma unroll
                for (int f = 0; f < 7; ++f) {
                    const double vl = lds.v[0][f][e];
                    phi[f] = J.hfx * (from_next_lane(vl) - vl);
                }
                double o[14];
                face_state_cot<METOD, true>(J, L, R, phi, wall_x, o);
                // the left state's cotangent stays with this cell (its plus
                // face), the right state's goes to the next cell (its minus face)
#pragma unroll
                for (int k = 0; k < 7; ++k) {
                    xdqp[k] = face_x ? o[k] : 0.0;
                    xdqm[k] = from_prev_lane(face_x ? o[7 + k] : 0.0);
                }
            }
            asm volatile("" ::: "memory");
            PrimM qw, q, qe;
            row_window(p, qw, q, qe);
            double xown[7], xtp[7], xtn[7], Gx[7];
            spread7<ORDER>(qw, q, qe, xdqp, xdqm, xown, xtp, xtn);
#pragma unroll
            for (int k = 0; k < 7; ++k) Gx[k] = (from_prev_lane(xtn[k]) + xown[k]) + from_next_lane(xtp[k]);
            double* gc = J.cx + (long long)p * nx;
#pragma unroll
            for (int k = 0; k < 7; ++k) st_buf(buf_rsrc(uniform_ptr(gc + k * P)), oo, Gx[k]);
            if (gslot >= 0) {
                double* g = J.gx + ((long long)gslot * ny + p) * 7;
#pragma unroll
                for (int k = 0; k < 7; ++k) g[k] = Gx[k];
            }
        }
    }
}

// K(q) . V of the Powell term of a cell (its derivative with respect to the
// cell's divergence estimate), times -1/dV.
__device__ __forceinline__ double powell_k(const MhdAdjArgs& J, const PrimM& q, long long k) {
    const long long P = J.plane;
    const double vmx = J.V[P + k], vmy = J.V[2 * P + k], vbx = J.V[3 * P + k], vby = J.V[4 * P + k],
                 ve = J.V[6 * P + k];
    const double K = (vmx * q.bx + vmy * q.by + ve * (q.u * q.bx + q.v * q.by)) * (1.0 / kCpi4) + vbx * q.u + vby * q.v;
    return -K * J.idV;
}

// primitive -> conserved cotangent at a cell:
// u = mx/r, v = my/r, p = g1 (E - (mx u + my v)/2 - |B|^2 / 8pi)
__device__ __forceinline__ void prim_to_cons_cot(const MhdAdjArgs& J, const PrimM& q, const double (&Pq)[7],
                                                 double (&T)[7]) {
    const double ir = fast_rcp(q.r);
    const double gp = J.g1 * Pq[3];
    T[0] = Pq[0] - (q.u * Pq[1] + q.v * Pq[2]) * ir + gp * 0.5 * (q.u * q.u + q.v * q.v);
    T[1] = Pq[1] * ir - gp * q.u;
    T[2] = Pq[2] * ir - gp * q.v;
    T[3] = Pq[4] - gp * q.bx * (1.0 / kCpi4);
    T[4] = Pq[5] - gp * q.by * (1.0 / kCpi4);
    T[5] = Pq[6] - gp * q.bz * (1.0 / kCpi4);
    T[6] = gp;
}

constexpr int kGatherW = 60;  // output columns per wave of the gather march

// The gather march. Lane l holds column c = cbase - 2 + l (cbase = 60 tx);
// lanes 2..61 own the strip's cells. Per row every lane loads its cell
// (through the boundary map) and the two y-face-state cotangents it owns
// (left state of face j+1/2, right state of face j-1/2) and spreads them
// along its column; the shares are carried, so row j is finished once the
// spread of row j+1 exists. A cell's cotangent is its x-gather (from the face
// march) + its y-gather + the Powell terms, mapped to conserved variables.
// The strip at the bottom (top) of the grid also marches the ghost rows -2,
// -1 (ny, ny+1): their y-gathers go to the side buffer gy.
template <int ORDER>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void mhd_vjp_gather_kernel(const MhdAdjArgs J) {
    const int lane = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % J.tiles_x, ty = tile / J.tiles_x;
    const int nx = J.nx, ny = J.ny;
    const int cbase = tx * kGatherW;
    const int c = cbase - 2 + lane;
    const int j0 = ty * J.rows, j1 = min(j0 + J.rows, ny);
    const int ja = j0 == 0 ? -2 : j0;       // first row whose gather this wave forms
    const int jb = j1 == ny ? ny + 1 : j1 - 1;  // last
    const bool col_in = c >= 0 && c < nx;
    const bool out_lane = col_in && lane >= 2 && lane <= 61;
    bool fx;
    const int ci = map_index(c, nx, J.bc[0], J.bc[1], fx);
    const unsigned oc = (unsigned)ci * 8u;
    const unsigned ocol = col_in ? (unsigned)c * 8u : kOffNone;  // y-faces, x-gather and V of this column
    const long long P = J.plane;
    const double hx = 0.5 * J.dy, hy = 0.5 * J.dx;

    auto load_q_raw = [&](int jl, double (&w)[7]) -> bool {
        bool fy;
        const int jj = map_index(jl, ny, J.bc[2], J.bc[3], fy);
        const double* b = J.U + (long long)jj * nx;
#pragma unroll
        for (int f = 0; f < 7; ++f) w[f] = ld_buf(buf_rsrc(uniform_ptr(b + f * P)), oc);
        return fy;
    };
    auto to_prim_q = [&](const double (&w)[7], bool fy) -> PrimM {
        PrimM q = prim_of_cons7(w, J.g1);
        if (fx) {
            q.u = -q.u;
            q.bx = -q.bx;
        }
        if (fy) {
            q.v = -q.v;
            q.by = -q.by;
        }
        return q;
    };
    auto load_q = [&](int jl) -> PrimM {
        double w[7];
        const bool fy = load_q_raw(jl, w);
        return to_prim_q(w, fy);
    };
    // y-face-state cotangents of the cell of row r: left state of face r+1/2
    // (face row r+1), right state of face r-1/2 (face row r); 0 outside
    auto load_yfaces = [&](int r, double (&dqp)[7], double (&dqm)[7]) {
        const bool hp = r + 1 >= 0 && r + 1 <= ny, hm = r >= 0 && r <= ny;
        const double* wp = J.wy + (long long)(hp ? r + 1 : 0) * nx;
        const double* wm = J.wy + (long long)(hm ? r : 0) * nx;
        const unsigned op = hp ? ocol : kOffNone, om = hm ? ocol : kOffNone;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            dqp[k] = ld_buf(buf_rsrc(uniform_ptr(wp + k * J.nfy)), op);
            dqm[k] = ld_buf(buf_rsrc(uniform_ptr(wm + (7 + k) * J.nfy)), om);
        }
    };
    // K . V of the Powell term of this lane's cell in row r (powell_k), 0
    // outside the grid; V of the cell in vv (mx, my, bx, by, E components)
    auto load_v_row = [&](int r, double (&vv)[5]) {
        const bool in = r >= 0 && r < ny;
        const double* b = J.V + (long long)(in ? r : 0) * nx;
        const unsigned o = in ? ocol : kOffNone;
        vv[0] = ld_buf(buf_rsrc(uniform_ptr(b + P)), o);
        vv[1] = ld_buf(buf_rsrc(uniform_ptr(b + 2 * P)), o);
        vv[2] = ld_buf(buf_rsrc(uniform_ptr(b + 3 * P)), o);
        vv[3] = ld_buf(buf_rsrc(uniform_ptr(b + 4 * P)), o);
        vv[4] = ld_buf(buf_rsrc(uniform_ptr(b + 6 * P)), o);
    };
    auto powell_k_row = [&](int r, const PrimM& q, const double (&vv)[5]) -> double {
        const bool in = r >= 0 && r < ny;
        const double K = (vv[0] * q.bx + vv[1] * q.by + vv[4] * (q.u * q.bx + q.v * q.by)) * (1.0 / kCpi4) +
                         vv[2] * q.u + vv[3] * q.v;
        return (in && col_in) ? -K * J.idV : 0.0;
    };
    auto powell_row = [&](int r, const PrimM& q, double (&vv)[5]) -> double {
        load_v_row(r, vv);
        return powell_k_row(r, q, vv);
    };

    // prologue: rows ja-2 .. ja+1; y-spreads of rows ja-1 (its share to row
    // ja) and ja
    PrimM qm = load_q(ja - 1), q0 = load_q(ja), qp = load_q(ja + 1);
    double tn_prev[7], yown[7], ytn[7];
    {
        const PrimM qmm = load_q(ja - 2);
        double dqp[7], dqm[7], own[7], tp[7];
        load_yfaces(ja - 1, dqp, dqm);
        spread7<ORDER>(qmm, qm, q0, dqp, dqm, own, tp, tn_prev);
        load_yfaces(ja, dqp, dqm);
        spread7<ORDER>(qm, q0, qp, dqp, dqm, yown, tp, ytn);
    }
    double v0[5], vtmp[5];
    double Km = powell_row(ja - 1, qm, vtmp);
    double K0 = powell_row(ja, q0, v0);

    for (int j = ja; j <= jb; ++j) {
        // every load of the pass is issued here, before any of them is
        // used: row j+2's state, row j+1's cotangent and y-face-state
        // cotangents (used first), then the cell's x-gather and the affine
        // terms of row j (0 off the grid's rows and owned cells), still in
        // flight while row j+1 is spread
        double w2[7], vp[5], ydqp[7], ydqm[7];
        const bool fy2 = load_q_raw(j + 2, w2);
        load_v_row(j + 1, vp);
        load_yfaces(j + 1, ydqp, ydqm);
        const bool in_row = j >= 0 && j < ny;
        const unsigned orow = in_row ? ocol : kOffNone;
        const unsigned oown = (in_row && out_lane) ? (unsigned)c * 8u : kOffNone;
        const long long kr = (long long)(in_row ? j : 0) * nx;
        // absent affine terms: the loads are still issued (no branch
        // between the loads), with an offset that reads nothing
        const double* Ab = J.A ? J.A : J.cx;
        const double* Bb = J.B ? J.B : J.cx;
        const unsigned oa = J.A ? oown : kOffNone, ob = J.B ? oown : kOffNone;
        double Gx[7], Av[7], Bv[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            Gx[k] = ld_buf(buf_rsrc(uniform_ptr(J.cx + kr + k * P)), orow);
            Av[k] = ld_buf(buf_rsrc(uniform_ptr(Ab + kr + k * P)), oa);
            Bv[k] = ld_buf(buf_rsrc(uniform_ptr(Bb + kr + k * P)), ob);
        }
        // row j+1's y-spread (rows j .. j+2) and Powell product
        const PrimM qpp = to_prim_q(w2, fy2);
        const double Kp = powell_k_row(j + 1, qp, vp);
        double own1[7], tp1[7], tn1[7];
        spread7<ORDER>(q0, qp, qpp, ydqp, ydqm, own1, tp1, tn1);
        double Gy[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) Gy[k] = (tn_prev[k] + yown[k]) + tp1[k];

        if (j >= 0 && j < ny) {
            // x: the cell's x-gather (face march, in Gx) and the neighbours'
            // bx and Powell products by lane shifts
            const double bxw = from_prev_lane(q0.bx), bxe = from_next_lane(q0.bx);
            const double Kw = from_prev_lane(K0), Ke = from_next_lane(K0);
            if (out_lane) {
                double Pq[7];
#pragma unroll
                for (int k = 0; k < 7; ++k) Pq[k] = Gx[k] + Gy[k];
                // Powell term of this cell: -(1/dV) Pdiv K(q) with
                // K = (0, bx/4pi, by/4pi, u, v, 0, (u bx + v by)/4pi) and
                // Pdiv = dy (bx_e - bx_w)/2 + dx (by_n - by_s)/2 of the
                // neighbours (ghosts through the boundary map)
                const PrimM& q = q0;
                const double pdiv = J.dy * (bxe - bxw) * 0.5 + J.dx * (qp.by - qm.by) * 0.5;
                {
                    const double vmx = v0[0], vmy = v0[1], vbx = v0[2], vby = v0[3], ve = v0[4];
                    const double sc = -pdiv * J.idV;
                    Pq[4] += sc * (vmx + ve * q.u) * (1.0 / kCpi4);
                    Pq[5] += sc * (vmy + ve * q.v) * (1.0 / kCpi4);
                    Pq[1] += sc * (vbx + ve * q.bx * (1.0 / kCpi4));
                    Pq[2] += sc * (vby + ve * q.by * (1.0 / kCpi4));
                }
                // ... and the Powell terms of the neighbours in the grid whose
                // divergence estimate reads this cell's bx (west / east) or
                // by (south / north): K is 0 outside the grid
                Pq[4] -= hx * Ke;
                Pq[4] += hx * Kw;
                Pq[5] -= hy * Kp;
                Pq[5] += hy * Km;
                double T[7];
                prim_to_cons_cot(J, q, Pq, T);
#pragma unroll
                for (int f = 0; f < 7; ++f) {
                    double o = J.c * T[f];
                    if (J.A) o += J.a * Av[f];
                    if (J.B) o += J.b * Bv[f];
                    st_buf(buf_rsrc(uniform_ptr(J.W + kr + f * P)), oown, o);
                }
            }
        } else if (col_in) {
            // a ghost row: its y-gather goes to the side buffer
            const int gr = j == -2 ? 0 : (j == -1 ? 1 : (j == ny ? 2 : 3));
            double* g = J.gy + ((long long)gr * nx + c) * 7;
#pragma unroll
            for (int k = 0; k < 7; ++k) g[k] = Gy[k];
        }

#pragma unroll
        for (int k = 0; k < 7; ++k) {
            tn_prev[k] = ytn[k];
            yown[k] = own1[k];
            ytn[k] = tn1[k];
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) v0[k] = vp[k];
        qm = q0;
        q0 = qp;
        qp = qpp;
        Km = K0;
        K0 = Kp;
    }
}

// The gathers of the ghost columns and rows added to the cells they copy
// (mirrored components negated), with the Powell terms of the edge cells
// whose divergence estimate reads a ghost that copies the cell: one thread
// per cell of the boundary band (columns 0, 1, nx-2, nx-1 and rows 0, 1,
// ny-2, ny-1), W += c * (primitive -> conserved)(those cotangents).
__global__ __launch_bounds__(64) void mhd_vjp_fold_kernel(const MhdAdjArgs J) {
    const int nx = J.nx, ny = J.ny;
    const double hx = 0.5 * J.dy, hy = 0.5 * J.dx;
    // thread t < 4 ny: column slot t % 4 (0, 1, nx-2, nx-1) of row t / 4;
    // then rows 0, 1, ny-2, ny-1 of the columns 2 .. nx-3
    const long long n1 = 4LL * ny, n2 = 4LL * nx;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n1 + n2;
         t += (long long)gridDim.x * blockDim.x) {
        int i, j;
        if (t < n1) {
            const int a = (int)(t & 3);
            j = (int)(t >> 2);
            i = a < 2 ? a : nx - 4 + a;
            // slots that repeat an earlier one (narrow grids) or leave the grid
            // (nx-2 and nx-1 repeat column 0 or 1 exactly when they are below 2)
            if (i < 0 || i >= nx || (a >= 2 && i < 2)) continue;
        } else {
            const long long t2 = t - n1;
            i = (int)(t2 % nx);
            const int b = (int)(t2 / nx);
            j = b < 2 ? b : ny - 4 + b;
            if (i < 2 || i >= nx - 2) continue;  // a band column: done above
            if (j < 0 || j >= ny || (b >= 2 && j < 2)) continue;
        }
        // which ghosts copy this cell (x: columns -2, -1, nx, nx+1; y: rows
        // -2, -1, ny, ny+1) and with which flip; every load below is issued
        // unconditionally from a valid address, then selected
        bool mx[4], fx4[4], my[4], fy4[4];
        bool any = false;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            mx[g] = map_index(g < 2 ? g - 2 : nx + g - 2, nx, J.bc[0], J.bc[1], fx4[g]) == i;
            my[g] = map_index(g < 2 ? g - 2 : ny + g - 2, ny, J.bc[2], J.bc[3], fy4[g]) == j;
            any = any || mx[g] || my[g];
        }
        if (!any) continue;  // no ghost copies this cell
        double Pq[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const double* Gx = J.gx + ((long long)g * ny + j) * 7;
            const double* Gy = J.gy + ((long long)g * nx + i) * 7;
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const double ax = Gx[k], ay = Gy[k];
                Pq[k] += mx[g] ? ((fx4[g] && (k == 1 || k == 4)) ? -ax : ax) : 0.0;
                Pq[k] += my[g] ? ((fy4[g] && (k == 2 || k == 5)) ? -ay : ay) : 0.0;
            }
        }
        // the Powell terms of the edge cells whose west / east / south /
        // north ghost copies this cell (mx[1], mx[2], my[1], my[2])
        {
            const long long kw = (long long)j * nx, ke = kw + nx - 1, ks = i, kn = (long long)(ny - 1) * nx + i;
            const double Kw = powell_k(J, adj_cell_m(J, 0, j), kw), Ke = powell_k(J, adj_cell_m(J, nx - 1, j), ke);
            const double Ks = powell_k(J, adj_cell_m(J, i, 0), ks), Kn = powell_k(J, adj_cell_m(J, i, ny - 1), kn);
            if (mx[1]) Pq[4] -= (fx4[1] ? -hx : hx) * Kw;
            if (mx[2]) Pq[4] += (fx4[2] ? -hx : hx) * Ke;
            if (my[1]) Pq[5] -= (fy4[1] ? -hy : hy) * Ks;
            if (my[2]) Pq[5] += (fy4[2] ? -hy : hy) * Kn;
        }
        const PrimM q = adj_cell_m(J, i, j);
        double T[7];
        prim_to_cons_cot(J, q, Pq, T);
        const long long k = (long long)j * nx + i;
#pragma unroll
        for (int f = 0; f < 7; ++f) J.W[f * J.plane + k] += J.c * T[f];
    }
}

}  // namespace gd2
