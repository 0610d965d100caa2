// gd2_mhd_adjoint.hpp -- discrete adjoint (vector-Jacobian product) of the
// ideal-MHD stage operator of gd2_mhd.hpp (HLLC-MHD or HLLD face flux,
// minmod MUSCL on the 7 primitives, Powell source) on gfx950.
//
// For L(U) = -(1/dV) (sum of face fluxes x face lengths + Powell term) the
// kernels compute W = a*A + b*B + c * J_L(U)^T V in two passes:
//
//  1. one thread per face (x-faces, then y-faces): the face cotangent
//     phi = (h/dV)(V_right - V_left) (zero on the side of a ghost cell) is
//     taken through the face flux to its two reconstructed states. The flux
//     Jacobian is formed by forward-mode tangents, one per reconstructed
//     input (7 per side), through a restatement of hlld_face over tangent
//     numbers that follows the same branches; each product phi . dF/dx_k goes
//     to plane k of a face buffer (planes 0..6: left state, 7..13: right).
//     Reflective west / south wall faces are evaluated mirrored, as the
//     forward stage does.
//  2. one thread per cell: gathers the face-state cotangents through the
//     minmod reconstruction of itself and its two neighbours on each line
//     (the branch the forward takes, 1/2-1/2 on ties, the rho / p <= 0 face
//     fall-back as a copy), adds the gathers of the boundary ghost cells that
//     copy it (mirrored components negated), adds the derivatives of the
//     Powell term, maps primitive to conserved cotangents and forms W.
//
// The forward stage's clamps are not differentiated (they do not fire at the
// smooth states an adjoint is linearised at).
#pragma once

#include "gd2_adjoint.hpp"
#include "gd2_mhd.hpp"

namespace gd2 {

// Forward-mode tangent number: value and derivative along one direction.
struct Tg {
    double v, d;
};

__device__ __forceinline__ Tg operator+(Tg a, Tg b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Tg operator-(Tg a, Tg b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Tg operator-(Tg a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Tg operator*(Tg a, Tg b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Tg operator*(double s, Tg a) { return {s * a.v, s * a.d}; }
__device__ __forceinline__ Tg operator*(Tg a, double s) { return {a.v * s, a.d * s}; }
__device__ __forceinline__ Tg operator+(Tg a, double s) { return {a.v + s, a.d}; }
__device__ __forceinline__ Tg operator-(Tg a, double s) { return {a.v - s, a.d}; }
__device__ __forceinline__ Tg trcp(Tg a) {
    const double r = 1.0 / a.v;
    return {r, -a.d * r * r};
}
__device__ __forceinline__ Tg tsqrt(Tg a) {
    const double s = sqrt(a.v);
    return {s, s > 0.0 ? 0.5 * a.d / s : 0.0};
}
__device__ __forceinline__ Tg tabs(Tg a) { return a.v < 0.0 ? -a : a; }
__device__ __forceinline__ Tg tmin(Tg a, Tg b) { return a.v <= b.v ? a : b; }
__device__ __forceinline__ Tg tmax(Tg a, Tg b) { return a.v >= b.v ? a : b; }

// hlld_face over tangent numbers: same formulas and branch order, IEEE
// division and square root (the branch tests see values within rounding of
// the forward's).
template <int METOD>
__device__ __forceinline__ void hlld_face_tg(const Tg (&in)[16], double gamma, double inv_g1, Tg (&P)[8]) {
    const Tg ro_L = in[0], p_L = in[1], u1 = in[2], v1 = in[3], w1 = in[4], bn1 = in[5], bt1 = in[6],
             bm1 = in[7];
    const Tg ro_R = in[8], p_R = in[9], u2 = in[10], v2 = in[11], w2 = in[12], bn2 = in[13], bt2 = in[14],
             bm2 = in[15];
    const Tg bb_L = bn1 * bn1 + bt1 * bt1 + bm1 * bm1;
    const Tg bb_R = bn2 * bn2 + bt2 * bt2 + bm2 * bm2;
    const Tg uu_L = 0.5 * (u1 * u1 + v1 * v1 + w1 * w1);
    const Tg uu_R = 0.5 * (u2 * u2 + v2 * v2 + w2 * w2);
    const Tg gpL = gamma * p_L + bb_L, gpR = gamma * p_R + bb_R;
    const Tg iroL = trcp(ro_L), iroR = trcp(ro_R);
    const Tg cfL = tsqrt((gpL + tsqrt(gpL * gpL - (4.0 * gamma) * p_L * (bn1 * bn1))) * (0.5 * iroL));
    const Tg cfR = tsqrt((gpR + tsqrt(gpR * gpR - (4.0 * gamma) * p_R * (bn2 * bn2))) * (0.5 * iroR));
    const Tg cmax = tmax(cfL, cfR);
    const Tg SL = tmin(u1, u2) - cmax;
    const Tg SR = tmax(u1, u2) + cmax;
    const Tg smax = tmax(tabs(SL), tabs(SR));
    const Tg pTL = p_L + 0.5 * bb_L, pTR = p_R + 0.5 * bb_R;
    const Tg suR = SR - u2, suL = SL - u1;
    const Tg mR = suR * ro_R, mL = suL * ro_L;
    const Tg idm = trcp(mR - mL);
    const Tg SM = (mR * u2 - mL * u1 - pTR + pTL) * idm;
    const Tg e1 = p_L * inv_g1 + ro_L * uu_L + 0.5 * bb_L;
    const Tg e2 = p_R * inv_g1 + ro_R * uu_R + 0.5 * bb_R;

    Tg FL[8], FR[8], UL[8], UR[8];
    const Tg zero = {0.0, 0.0};
    FL[0] = ro_L * u1;
    FL[1] = ro_L * u1 * u1 + pTL - bn1 * bn1;
    FL[2] = ro_L * u1 * v1 - bn1 * bt1;
    FL[3] = ro_L * u1 * w1 - bn1 * bm1;
    FL[4] = (e1 + pTL) * u1 - bn1 * (u1 * bn1 + v1 * bt1 + w1 * bm1);
    FL[5] = zero;
    FL[6] = u1 * bt1 - v1 * bn1;
    FL[7] = u1 * bm1 - w1 * bn1;
    FR[0] = ro_R * u2;
    FR[1] = ro_R * u2 * u2 + pTR - bn2 * bn2;
    FR[2] = ro_R * u2 * v2 - bn2 * bt2;
    FR[3] = ro_R * u2 * w2 - bn2 * bm2;
    FR[4] = (e2 + pTR) * u2 - bn2 * (u2 * bn2 + v2 * bt2 + w2 * bm2);
    FR[5] = zero;
    FR[6] = u2 * bt2 - v2 * bn2;
    FR[7] = u2 * bm2 - w2 * bn2;
    UL[0] = ro_L; UL[1] = ro_L * u1; UL[2] = ro_L * v1; UL[3] = ro_L * w1;
    UL[4] = e1; UL[5] = bn1; UL[6] = bt1; UL[7] = bm1;
    UR[0] = ro_R; UR[1] = ro_R * u2; UR[2] = ro_R * v2; UR[3] = ro_R * w2;
    UR[4] = e2; UR[5] = bn2; UR[6] = bt2; UR[7] = bm2;

    const Tg iS = trcp(SR - SL);
    const Tg bn = (SR * bn2 - SL * bn1) * iS;
    const Tg bt = (SR * bt2 - SL * bt1 + FL[6] - FR[6]) * iS;
    const Tg bm = (SR * bm2 - SL * bm1 + FL[7] - FR[7]) * iS;
    const Tg bbn = bn * bn;
    const Tg iLM = trcp(SL - SM), iRM = trcp(SR - SM);
    const Tg ro_LL = ro_L * suL * iLM;
    const Tg ro_RR = ro_R * suR * iRM;

    if constexpr (METOD == 2) {
        const Tg sbv1 = u1 * bn1 + v1 * bt1 + w1 * bm1;
        const Tg sbv2 = u2 * bn2 + v2 * bt2 + w2 * bm2;
        const Tg UZ0 = (SR * UR[0] - SL * UL[0] + FL[0] - FR[0]) * iS;
        const Tg UZ1 = (SR * UR[1] - SL * UL[1] + FL[1] - FR[1]) * iS;
        const Tg UZ2 = (SR * UR[2] - SL * UL[2] + FL[2] - FR[2]) * iS;
        const Tg UZ3 = (SR * UR[3] - SL * UL[3] + FL[3] - FR[3]) * iS;
        const Tg suRm = suR * iRM, suLm = suL * iLM;
        const Tg ptz = 0.5 * ((pTR + ro_R * suR * (SM - u2)) + (pTL + ro_L * suL * (SM - u1)));
        const Tg iUZ0 = trcp(UZ0);
        const Tg vv = UZ2 * iUZ0, ww = UZ3 * iUZ0;
        const Tg sbvz = (bn * UZ1 + bt * UZ2 + bm * UZ3) * iUZ0;
        const Tg ee2 = e2 * suRm + (ptz * SM - pTR * u2 + bn * (sbv2 - sbvz)) * iRM;
        const Tg ee1 = e1 * suLm + (ptz * SM - pTL * u1 + bn * (sbv1 - sbvz)) * iLM;
        const bool left = (SL.v >= 0.0) || (SM.v >= 0.0);
        const Tg S = (SL.v >= 0.0) ? zero : (left ? SL : (SR.v > 0.0 ? SR : zero));
        const Tg rs = left ? ro_LL : ro_RR;
        const Tg es = left ? ee1 : ee2;
        const Tg Us[8] = {rs, rs * SM, rs * vv, rs * ww, es, bn, bt, bm};
#pragma unroll
        for (int i = 0; i < 8; ++i) P[i] = (left ? FL[i] : FR[i]) + S * Us[i] - S * (left ? UL[i] : UR[i]);
    } else {
        const Tg PTT = (mR * pTL - mL * pTR + ro_L * ro_R * suR * suL * (u2 - u1)) * idm;
        const Tg ttL = ro_L * suL * (SL - SM) - bbn;
        const Tg ttR = ro_R * suR * (SR - SM) - bbn;
        Tg vLL, wLL, vRR, wRR, btt1, bmm1, btt2, bmm2;
        if (fabs(ttL.v) >= 0.00001) {
            const Tg it = trcp(ttL);
            vLL = v1 - bn * bt1 * (SM - u1) * it;
            wLL = w1 - bn * bm1 * (SM - u1) * it;
            const Tg q = (ro_L * suL * suL - bbn) * it;
            btt1 = bt1 * q;
            bmm1 = bm1 * q;
        } else {
            vLL = v1; wLL = w1; btt1 = zero; bmm1 = zero;
        }
        if (fabs(ttR.v) >= 0.00001) {
            const Tg it = trcp(ttR);
            vRR = v2 - bn * bt2 * (SM - u2) * it;
            wRR = w2 - bn * bm2 * (SM - u2) * it;
            const Tg q = (ro_R * suR * suR - bbn) * it;
            btt2 = bt2 * q;
            bmm2 = bm2 * q;
        } else {
            vRR = v2; wRR = w2; btt2 = zero; bmm2 = zero;
        }
        const Tg eLL = (e1 * suL + PTT * SM - pTL * u1 +
                        bn * ((u1 * bn1 + v1 * bt1 + w1 * bm1) - (SM * bn + vLL * btt1 + wLL * bmm1))) * iLM;
        const Tg eRR = (e2 * suR + PTT * SM - pTR * u2 +
                        bn * ((u2 * bn2 + v2 * bt2 + w2 * bm2) - (SM * bn + vRR * btt2 + wRR * bmm2))) * iRM;
        const Tg sLL = tsqrt(ro_LL), sRR = tsqrt(ro_RR);
        Tg SLL = SM - tabs(bn) * trcp(sLL);
        Tg SRR = SM + tabs(bn) * trcp(sRR);
        double idbn;
        if (fabs(bn.v) > 0.000001) {
            idbn = bn.v > 0.0 ? 1.0 : -1.0;
        } else {
            idbn = 0.0;
            SLL = SM;
            SRR = SM;
        }
        const Tg isum = trcp(sLL + sRR);
        const Tg vLLL = (sLL * vLL + sRR * vRR + idbn * (btt2 - btt1)) * isum;
        const Tg wLLL = (sLL * wLL + sRR * wRR + idbn * (bmm2 - bmm1)) * isum;
        const Tg bttt = (sLL * btt2 + sRR * btt1 + idbn * sLL * sRR * (vRR - vLL)) * isum;
        const Tg bmmm = (sLL * bmm2 + sRR * bmm1 + idbn * sLL * sRR * (wRR - wLL)) * isum;
        const Tg vb3 = SM * bn + vLLL * bttt + wLLL * bmmm;
        const Tg eLLL = eLL - idbn * sLL * ((SM * bn + vLL * btt1 + wLL * bmm1) - vb3);
        const Tg eRRR = eRR + idbn * sRR * ((SM * bn + vRR * btt2 + wRR * bmm2) - vb3);
        const bool left = SM.v >= 0.0;
        const Tg S1 = left ? SL : SR;
        const Tg S2 = left ? SLL : SRR;
        const Tg rs = left ? ro_LL : ro_RR;
        const Tg vs = left ? vLL : vRR, ws = left ? wLL : wRR;
        const Tg bts = left ? btt1 : btt2, bms = left ? bmm1 : bmm2;
        const Tg es = left ? eLL : eRR;
        const Tg ess = left ? eLLL : eRRR;
        const bool outer = left ? (SL.v >= 0.0) : !(SRR.v > 0.0 || SR.v > 0.0);
        const bool single = left ? (SLL.v >= 0.0) : !(SRR.v > 0.0);
        const Tg Us[8] = {rs, rs * SM, rs * vs, rs * ws, es, bn, bts, bms};
        const Tg Uss[8] = {rs, rs * SM, rs * vLLL, rs * wLLL, ess, bn, bttt, bmmm};
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const Tg Fi = left ? FL[i] : FR[i];
            const Tg Ui = left ? UL[i] : UR[i];
            if (outer)
                P[i] = Fi;
            else if (single)
                P[i] = Fi + S1 * Us[i] - S1 * Ui;
            else
                P[i] = Fi + S2 * Uss[i] - (S2 - S1) * Us[i] - S1 * Ui;
        }
    }
    P[5] = -(smax * (bn2 - bn1));
}

// mhd_flux over tangent numbers. L, R: r, u, v, p, bx, by, bz (Gaussian B);
// F: lab-frame fluxes in conserved order r, mx, my, bx, by, bz, e.
template <int METOD, bool XFACE>
__device__ __forceinline__ void mhd_flux_tg(const Tg (&L)[7], const Tg (&R)[7], double gamma, double inv_g1,
                                            Tg (&F)[7]) {
    const double is = 1.0 / kSpi4;
    const Tg z = {0.0, 0.0};
    Tg in[16];
    if constexpr (XFACE) {
        const Tg a[16] = {L[0], L[3], L[1], L[2], z, L[4] * is, L[5] * is, L[6] * is,
                          R[0], R[3], R[1], R[2], z, R[4] * is, R[5] * is, R[6] * is};
#pragma unroll
        for (int i = 0; i < 16; ++i) in[i] = a[i];
    } else {
        const Tg a[16] = {L[0], L[3], L[2], z, L[1], L[5] * is, L[6] * is, L[4] * is,
                          R[0], R[3], R[2], z, R[1], R[5] * is, R[6] * is, R[4] * is};
#pragma unroll
        for (int i = 0; i < 16; ++i) in[i] = a[i];
    }
    Tg P[8];
    hlld_face_tg<METOD>(in, gamma, inv_g1, P);
    F[0] = P[0];
    F[6] = P[4];
    if constexpr (XFACE) {
        F[1] = P[1];
        F[2] = P[2];
        F[3] = kSpi4 * P[5];
        F[4] = kSpi4 * P[6];
        F[5] = kSpi4 * P[7];
    } else {
        F[2] = P[1];
        F[1] = P[3];
        F[4] = kSpi4 * P[5];
        F[5] = kSpi4 * P[6];
        F[3] = kSpi4 * P[7];
    }
}

struct MhdAdjArgs {
    const double* U;  // state the stage operator is linearised at (7 SoA planes)
    const double* V;  // cotangent
    const double* A;  // optional
    const double* B;  // optional
    double* W;        // output
    double a, b, c;   // W = a*A + b*B + c*J^T V
    double* wx;       // x-face state cotangents: 14 planes of (nx+1) x ny
    double* wy;       // y-face state cotangents: 14 planes of nx x (ny+1)
    int nx, ny;
    long long plane;
    long long nfx, nfy;
    double dx, dy, dV;
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

// One thread per face f + 1/2 of a line (f = -1 .. n-1): the cotangents of
// its reconstructed left and right states.
template <int ORDER, int METOD, bool XFACE>
__global__ __launch_bounds__(256) void mhd_vjp_face_kernel(const MhdAdjArgs J) {
    const long long nf = XFACE ? J.nfx : J.nfy;
    double* out = XFACE ? J.wx : J.wy;
    const int n = XFACE ? J.nx : J.ny;
    const double h = (XFACE ? J.dy : J.dx) / J.dV;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nf;
         t += (long long)gridDim.x * blockDim.x) {
        int i, j, f;  // the face lies between cell f and cell f+1 of its line
        if constexpr (XFACE) {
            f = (int)(t % (J.nx + 1)) - 1;
            j = (int)(t / (J.nx + 1));
            i = f;
        } else {
            i = (int)(t % J.nx);
            f = (int)(t / J.nx) - 1;
            j = f;
        }
        auto cell = [&](int o) { return XFACE ? adj_cell_m(J, i + o, j) : adj_cell_m(J, i, j + o); };
        const PrimM q0 = cell(-1), q1 = cell(0), q2 = cell(1), q3 = cell(2);
        const PrimM s1 = slope3m<ORDER>(q0, q1, q2), s2 = slope3m<ORDER>(q1, q2, q3);
        const PrimM L = face_m(q1, s1, 0.5), R = face_m(q2, s2, -0.5);

        // face cotangent: the face flux leaves cell f and enters cell f+1
        double phi[7];
        const long long kl = (long long)j * J.nx + i;
        const long long kr = XFACE ? kl + 1 : kl + J.nx;
#pragma unroll
        for (int c = 0; c < 7; ++c) {
            const double vl = f >= 0 ? J.V[c * J.plane + kl] : 0.0;
            const double vr = f + 1 < n ? J.V[c * J.plane + kr] : 0.0;
            phi[c] = h * (vr - vl);
        }
        // a reflective west / south wall face is evaluated mirrored:
        // F = unmirror(flux(mirror(R), mirror(L)))
        const bool wall = f == -1 && J.bc[XFACE ? 0 : 2] == GD2_BC_REFLECTIVE;
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
#pragma unroll 1
        for (int k = 0; k < 14; ++k) {
            Tg Xt[7], Yt[7];
#pragma unroll
            for (int c = 0; c < 7; ++c) {
                Xt[c] = Tg{comp_m(X, c), k == c ? 1.0 : 0.0};
                Yt[c] = Tg{comp_m(Y, c), k == c + 7 ? 1.0 : 0.0};
            }
            Tg F[7];
            mhd_flux_tg<METOD, XFACE>(Xt, Yt, J.gamma, J.inv_g1, F);
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < 7; ++c) s += phi[c] * F[c].d;
            int pl = k;
            if (wall) {
                const int comp = k < 7 ? k : k - 7;
                pl = k < 7 ? comp + 7 : comp;  // X is the mirrored right state
                if (comp == mu || comp == mb) s = -s;
            }
            out[pl * nf + t] = s;
        }
    }
}

// Primitive cotangent of cell e (extended index, -2 .. n+1) of a line (row
// `line` for x, column `line` for y) gathered from the face-state cotangents
// through its own reconstruction and those of cells e-1 and e+1.
template <int ORDER, bool XDIR>
__device__ __forceinline__ void gather_line_m(const MhdAdjArgs& J, int e, int line, double (&G)[7]) {
    const int n = XDIR ? J.nx : J.ny;
    const long long nf = XDIR ? J.nfx : J.nfy;
    const double* w = XDIR ? J.wx : J.wy;
    // plane c (left state) or 7 + c (right state) of face f + 1/2; 0 for
    // faces outside f = -1 .. n-1
    auto wface = [&](int f, int pl) -> double {
        if (f < -1 || f > n - 1) return 0.0;
        const long long id = XDIR ? (long long)line * (n + 1) + (f + 1) : (long long)(f + 1) * J.nx + line;
        return w[pl * nf + id];
    };
    if constexpr (ORDER == 1) {
#pragma unroll
        for (int c = 0; c < 7; ++c) G[c] = wface(e, c) + wface(e - 1, 7 + c);
        return;
    } else {
        PrimM q[5];
#pragma unroll
        for (int o = 0; o < 5; ++o) q[o] = XDIR ? adj_cell_m(J, e - 2 + o, line) : adj_cell_m(J, line, e - 2 + o);
#pragma unroll
        for (int c = 0; c < 7; ++c) {
            double acc = 0.0;
            // cell e-1+m (m = 0, 1, 2) spreads the cotangents of its plus value
            // (left state of face e-1+m+1/2) and minus value (right state of
            // face e-1+m-1/2) over its stencil; cell e takes its share
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const int cc = e - 1 + m;
                const double dqp = wface(cc, c), dqm = wface(cc - 1, 7 + c);
                const double qa = comp_m(q[m], c), qb = comp_m(q[m + 1], c), qd = comp_m(q[m + 2], c);
                const MMW wt = mm_weights(qb - qa, qd - qb);
                bool fp = false, fm = false;
                if (c == 0 || c == 3) {
                    const double s = minmod(qb - qa, qd - qb);
                    fp = !(qb + 0.5 * s > 0.0);
                    fm = !(qb + -0.5 * s > 0.0);
                }
                double own, to_prev, to_next;
                spread(dqp, dqm, wt, fp, fm, own, to_prev, to_next);
                acc += m == 0 ? to_next : (m == 1 ? own : to_prev);
            }
            G[c] = acc;
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
    return -K / J.dV;
}

template <int ORDER>
__global__ __launch_bounds__(256) void mhd_vjp_cell_kernel(const MhdAdjArgs J) {
    const long long ncell = J.plane;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % J.nx), j = (int)(t / J.nx);
        double Pq[7], G[7];
        gather_line_m<ORDER, true>(J, i, j, Pq);
        // ghost columns that copy column i
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int gc = g < 2 ? g - 2 : J.nx + g - 2;
            bool fl;
            if (map_index(gc, J.nx, J.bc[0], J.bc[1], fl) != i) continue;
            gather_line_m<ORDER, true>(J, gc, j, G);
            if (fl) {
                G[1] = -G[1];
                G[4] = -G[4];
            }
#pragma unroll
            for (int c = 0; c < 7; ++c) Pq[c] += G[c];
        }
        gather_line_m<ORDER, false>(J, j, i, G);
#pragma unroll
        for (int c = 0; c < 7; ++c) Pq[c] += G[c];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int gr = g < 2 ? g - 2 : J.ny + g - 2;
            bool fl;
            if (map_index(gr, J.ny, J.bc[2], J.bc[3], fl) != j) continue;
            gather_line_m<ORDER, false>(J, gr, i, G);
            if (fl) {
                G[2] = -G[2];
                G[5] = -G[5];
            }
#pragma unroll
            for (int c = 0; c < 7; ++c) Pq[c] += G[c];
        }

        // Powell term of this cell: -(1/dV) Pdiv K(q) with
        // K = (0, bx/4pi, by/4pi, u, v, 0, (u bx + v by)/4pi) and
        // Pdiv = dy (bx_e - bx_w)/2 + dx (by_n - by_s)/2 of the neighbours
        const PrimM q = adj_cell_m(J, i, j);
        const PrimM qw = adj_cell_m(J, i - 1, j), qe = adj_cell_m(J, i + 1, j);
        const PrimM qs = adj_cell_m(J, i, j - 1), qn = adj_cell_m(J, i, j + 1);
        const double pdiv = J.dy * (qe.bx - qw.bx) * 0.5 + J.dx * (qn.by - qs.by) * 0.5;
        {
            const long long P = J.plane;
            const double vmx = J.V[P + t], vmy = J.V[2 * P + t], vbx = J.V[3 * P + t], vby = J.V[4 * P + t],
                         ve = J.V[6 * P + t];
            const double s = -pdiv / J.dV;
            Pq[4] += s * (vmx + ve * q.u) * (1.0 / kCpi4);
            Pq[5] += s * (vmy + ve * q.v) * (1.0 / kCpi4);
            Pq[1] += s * (vbx + ve * q.bx * (1.0 / kCpi4));
            Pq[2] += s * (vby + ve * q.by * (1.0 / kCpi4));
        }
        // ... and the Powell terms of the cells whose divergence estimate
        // reads this cell's bx (as a west / east neighbour) or by (south /
        // north), ghosts included
        const double hx = 0.5 * J.dy, hy = 0.5 * J.dx;
        if (i + 1 < J.nx) Pq[4] -= hx * powell_k(J, qe, t + 1);
        if (i - 1 >= 0) Pq[4] += hx * powell_k(J, qw, t - 1);
        if (j + 1 < J.ny) Pq[5] -= hy * powell_k(J, qn, t + J.nx);
        if (j - 1 >= 0) Pq[5] += hy * powell_k(J, qs, t - J.nx);
        {
            bool fl;
            if (map_index(-1, J.nx, J.bc[0], J.bc[1], fl) == i) {
                const long long k0 = (long long)j * J.nx;
                Pq[4] -= (fl ? -hx : hx) * powell_k(J, adj_cell_m(J, 0, j), k0);
            }
            if (map_index(J.nx, J.nx, J.bc[0], J.bc[1], fl) == i) {
                const long long k1 = (long long)j * J.nx + J.nx - 1;
                Pq[4] += (fl ? -hx : hx) * powell_k(J, adj_cell_m(J, J.nx - 1, j), k1);
            }
            if (map_index(-1, J.ny, J.bc[2], J.bc[3], fl) == j) {
                Pq[5] -= (fl ? -hy : hy) * powell_k(J, adj_cell_m(J, i, 0), i);
            }
            if (map_index(J.ny, J.ny, J.bc[2], J.bc[3], fl) == j) {
                const long long k1 = (long long)(J.ny - 1) * J.nx + i;
                Pq[5] += (fl ? -hy : hy) * powell_k(J, adj_cell_m(J, i, J.ny - 1), k1);
            }
        }

        // primitive -> conserved cotangent at this cell:
        // u = mx/r, v = my/r, p = g1 (E - (mx u + my v)/2 - |B|^2 / 8pi)
        const double ir = 1.0 / q.r;
        const double gp = J.g1 * Pq[3];
        double T[7];
        T[0] = Pq[0] - (q.u * Pq[1] + q.v * Pq[2]) * ir + gp * 0.5 * (q.u * q.u + q.v * q.v);
        T[1] = Pq[1] * ir - gp * q.u;
        T[2] = Pq[2] * ir - gp * q.v;
        T[3] = Pq[4] - gp * q.bx * (1.0 / kCpi4);
        T[4] = Pq[5] - gp * q.by * (1.0 / kCpi4);
        T[5] = Pq[6] - gp * q.bz * (1.0 / kCpi4);
        T[6] = gp;
        const long long P = J.plane;
#pragma unroll
        for (int c = 0; c < 7; ++c) {
            double o = J.c * T[c];
            if (J.A) o += J.a * J.A[c * P + t];
            if (J.B) o += J.b * J.B[c * P + t];
            J.W[c * P + t] = o;
        }
    }
}

}  // namespace gd2
