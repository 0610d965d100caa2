// gd2_mhd.hpp -- fused ideal-MHD stage kernel for gfx950 (vz = 0, 7 fields,
// Gaussian B as in the reference), HLLC-MHD or HLLD face flux, minmod MUSCL on
// the 7 primitives, Powell 8-wave source, conservative update.
//
// Same tiling as the Euler kernel (gd2_euler.hpp): one wave per 63-column
// strip marching up its rows, y-faces carried in registers, x-faces from a
// 3-row LDS ring, each face flux computed once.
#pragma once

#include "gd2_euler.hpp"
#include "gd2_host.hpp"

namespace gd2 {

// Primitive MHD state, Gaussian B.
struct PrimM {
    double r, u, v, p, bx, by, bz;
};

// HLLDQ_Korolkov (Gas_Dinamic.cu:964-1510) in its face frame, metod 2
// (HLLC-MHD) or 3 (HLLD). Inputs: density, pressure, velocity (n, t, m) and
// Heaviside-Lorentz field (n, t, m) on both sides. Output PO in the face
// frame: mass, momentum (n, t, m), energy, field flux (n, t, m) with the
// normal-field flux -max(|SL|,|SR|)(bn_R - bn_L) (Gas_Dinamic.cu:1488-1490).
struct FluxM {
    double f[8];  // 0 mass, 1..3 momentum n/t/m, 4 energy, 5..7 field n/t/m
};

template <int METOD>
__device__ __forceinline__ FluxM hlld_face(double ro_L, double p_L, double u1, double v1, double w1, double bn1,
                                           double bt1, double bm1, double ro_R, double p_R, double u2, double v2,
                                           double w2, double bn2, double bt2, double bm2, double gamma,
                                           double inv_g1, double& smax) {
    // Gas_Dinamic.cu:1047-1078
    const double bb_L = bn1 * bn1 + bt1 * bt1 + bm1 * bm1;
    const double bb_R = bn2 * bn2 + bt2 * bt2 + bm2 * bm2;
    const double uu_L = 0.5 * (u1 * u1 + v1 * v1 + w1 * w1);
    const double uu_R = 0.5 * (u2 * u2 + v2 * v2 + w2 * w2);
    const double gpL = gamma * p_L + bb_L, gpR = gamma * p_R + bb_R;
    const double iroL = fast_rcp(ro_L), iroR = fast_rcp(ro_R);
    const double cfL = fast_sqrt((gpL + fast_sqrt(gpL * gpL - 4.0 * gamma * p_L * (bn1 * bn1))) * (0.5 * iroL));
    const double cfR = fast_sqrt((gpR + fast_sqrt(gpR * gpR - 4.0 * gamma * p_R * (bn2 * bn2))) * (0.5 * iroR));
    const double cmax = fmax(cfL, cfR);
    const double SL = fmin(u1, u2) - cmax;
    const double SR = fmax(u1, u2) + cmax;
    smax = fmax(fabs(SL), fabs(SR));
    const double pTL = p_L + 0.5 * bb_L, pTR = p_R + 0.5 * bb_R;
    const double suR = SR - u2, suL = SL - u1;
    const double mR = suR * ro_R, mL = suL * ro_L;
    const double idm = fast_rcp(mR - mL);
    // separately rounded products: mirror-image states give SM = 0 exactly,
    // as in the reference's evaluation order (Gas_Dinamic.cu:1071)
    const double smR = mR * u2;
    const double smL = mL * u1;
    const double SM = (smR - smL - pTR + pTL) * idm;

    // HLL averages of the field (Gas_Dinamic.cu:1127-1130); the tangential
    // field fluxes u bt - v bn, u bm - w bn of both sides
    const double iS = fast_rcp(SR - SL);
    const double bn = (SR * bn2 - SL * bn1) * iS;
    const double bt = (SR * bt2 - SL * bt1 + (u1 * bt1 - v1 * bn1) - (u2 * bt2 - v2 * bn2)) * iS;
    const double bm = (SR * bm2 - SL * bm1 + (u1 * bm1 - w1 * bn1) - (u2 * bm2 - w2 * bn2)) * iS;
    const double bbn = bn * bn;
    const double iLM = fast_rcp(SL - SM), iRM = fast_rcp(SR - SM);
    const double ro_LL = ro_L * suL * iLM;
    const double ro_RR = ro_R * suR * iRM;

    // The flux is F + (star-state terms) of one side K (left when the region
    // formula takes it); F, U and the star energies are formed for side K
    // only, from its selected inputs (the same expressions on the same values
    // as forming both sides and selecting).
    FluxM P;
    if constexpr (METOD == 2) {
        // HLLC-MHD with HLL-averaged tangential velocity (Gas_Dinamic.cu:1137-1301)
        const double e1 = p_L * inv_g1 + ro_L * uu_L + 0.5 * bb_L;
        const double e2 = p_R * inv_g1 + ro_R * uu_R + 0.5 * bb_R;
        const double UZ0 = (SR * ro_R - SL * ro_L + ro_L * u1 - ro_R * u2) * iS;
        const double UZ1 = (SR * (ro_R * u2) - SL * (ro_L * u1) + (ro_L * u1 * u1 + pTL - bn1 * bn1) -
                            (ro_R * u2 * u2 + pTR - bn2 * bn2)) * iS;
        const double UZ2 = (SR * (ro_R * v2) - SL * (ro_L * v1) + (ro_L * u1 * v1 - bn1 * bt1) -
                            (ro_R * u2 * v2 - bn2 * bt2)) * iS;
        const double UZ3 = (SR * (ro_R * w2) - SL * (ro_L * w1) + (ro_L * u1 * w1 - bn1 * bm1) -
                            (ro_R * u2 * w2 - bn2 * bm2)) * iS;
        const double ptz = 0.5 * ((pTR + ro_R * suR * (SM - u2)) + (pTL + ro_L * suL * (SM - u1)));
        const double iUZ0 = fast_rcp(UZ0);
        const double vv = UZ2 * iUZ0, ww = UZ3 * iUZ0;
        const double sbvz = (bn * UZ1 + bt * UZ2 + bm * UZ3) * iUZ0;
        // branch order of Gas_Dinamic.cu:1250-1277: SL >= 0, SM >= 0, SR > 0, else
        const bool left = (SL >= 0.0) || (SM >= 0.0);
        const double S = (SL >= 0.0) ? 0.0 : (left ? SL : (SR > 0.0 ? SR : 0.0));
        const double r = left ? ro_L : ro_R, u = left ? u1 : u2, v = left ? v1 : v2, w = left ? w1 : w2;
        const double bnK = left ? bn1 : bn2, btK = left ? bt1 : bt2, bmK = left ? bm1 : bm2;
        const double pT = left ? pTL : pTR, e = left ? e1 : e2;
        const double su = left ? suL : suR, iKM = left ? iLM : iRM;
        const double sbv = u * bnK + v * btK + w * bmK;
        const double rs = left ? ro_LL : ro_RR;
        const double es = e * (su * iKM) + (ptz * SM - pT * u + bn * (sbv - sbvz)) * iKM;
        const double F[8] = {r * u, r * u * u + pT - bnK * bnK, r * u * v - bnK * btK, r * u * w - bnK * bmK,
                             (e + pT) * u - bnK * sbv, 0.0, u * btK - v * bnK, u * bmK - w * bnK};
        const double U[8] = {r, r * u, r * v, r * w, e, bnK, btK, bmK};
        const double Us[8] = {rs, rs * SM, rs * vv, rs * ww, es, bn, bt, bm};
#pragma unroll
        for (int i = 0; i < 8; ++i) P.f[i] = F[i] + S * Us[i] - S * U[i];
    } else {
        // HLLD (Gas_Dinamic.cu:1303-1484)
        const double PTT = (mR * pTL - mL * pTR + ro_L * ro_R * suR * suL * (u2 - u1)) * idm;
        const double ttL = ro_L * suL * (SL - SM) - bbn;
        const double ttR = ro_R * suR * (SR - SM) - bbn;
        double vLL, wLL, vRR, wRR, btt1, bmm1, btt2, bmm2;
        if (fabs(ttL) >= 0.00001) {
            const double it = fast_rcp(ttL);
            vLL = v1 - bn * bt1 * (SM - u1) * it;
            wLL = w1 - bn * bm1 * (SM - u1) * it;
            const double q = (ro_L * suL * suL - bbn) * it;
            btt1 = bt1 * q;
            bmm1 = bm1 * q;
        } else {
            vLL = v1; wLL = w1; btt1 = 0.0; bmm1 = 0.0;
        }
        if (fabs(ttR) >= 0.00001) {
            const double it = fast_rcp(ttR);
            vRR = v2 - bn * bt2 * (SM - u2) * it;
            wRR = w2 - bn * bm2 * (SM - u2) * it;
            const double q = (ro_R * suR * suR - bbn) * it;
            btt2 = bt2 * q;
            bmm2 = bm2 * q;
        } else {
            vRR = v2; wRR = w2; btt2 = 0.0; bmm2 = 0.0;
        }
        const double sLL = fast_sqrt(ro_LL), sRR = fast_sqrt(ro_RR);
        double SLL = SM - fabs(bn) * fast_rcp(sLL);
        double SRR = SM + fabs(bn) * fast_rcp(sRR);
        double idbn;
        if (fabs(bn) > 0.000001) {
            idbn = bn > 0.0 ? 1.0 : -1.0;
        } else {
            idbn = 0.0;
            SLL = SM;
            SRR = SM;
        }
        const double isum = fast_rcp(sLL + sRR);
        const double vLLL = (sLL * vLL + sRR * vRR + idbn * (btt2 - btt1)) * isum;
        const double wLLL = (sLL * wLL + sRR * wRR + idbn * (bmm2 - bmm1)) * isum;
        const double bttt = (sLL * btt2 + sRR * btt1 + idbn * sLL * sRR * (vRR - vLL)) * isum;
        const double bmmm = (sLL * bmm2 + sRR * bmm1 + idbn * sLL * sRR * (wRR - wLL)) * isum;
        const double vb3 = SM * bn + vLLL * bttt + wLLL * bmmm;
        // six regions (Gas_Dinamic.cu:1434-1484): F + S*(U* - U) for the
        // single-star regions, F + S**U** - (S** - S)U* - S U for the double-star
        const bool left = SM >= 0.0;
        const double r = left ? ro_L : ro_R, u = left ? u1 : u2, v = left ? v1 : v2, w = left ? w1 : w2;
        const double pK = left ? p_L : p_R, uuK = left ? uu_L : uu_R, bbK = left ? bb_L : bb_R;
        const double bnK = left ? bn1 : bn2, btK = left ? bt1 : bt2, bmK = left ? bm1 : bm2;
        const double pT = left ? pTL : pTR, su = left ? suL : suR, iKM = left ? iLM : iRM;
        const double S1 = left ? SL : SR;
        const double S2 = left ? SLL : SRR;
        const double rs = left ? ro_LL : ro_RR;
        const double vs = left ? vLL : vRR, ws = left ? wLL : wRR;
        const double bts = left ? btt1 : btt2, bms = left ? bmm1 : bmm2;
        const double sK = left ? sLL : sRR;
        const double sgn = left ? -idbn : idbn;
        const double e = pK * inv_g1 + r * uuK + 0.5 * bbK;
        const double sbv = u * bnK + v * btK + w * bmK;
        const double vbK = SM * bn + vs * bts + ws * bms;
        const double es = (e * su + PTT * SM - pT * u + bn * (sbv - vbK)) * iKM;
        const double ess = es + sgn * sK * (vbK - vb3);
        // region selectors in the reference's branch order: left side
        // SL >= 0 / SLL >= 0 / double star; right side SRR > 0 (double star) /
        // SR > 0 (single star) / outer
        const bool outer = left ? (SL >= 0.0) : !(SRR > 0.0 || SR > 0.0);
        const bool single = left ? (SLL >= 0.0) : !(SRR > 0.0);
        const double F[8] = {r * u, r * u * u + pT - bnK * bnK, r * u * v - bnK * btK, r * u * w - bnK * bmK,
                             (e + pT) * u - bnK * sbv, 0.0, u * btK - v * bnK, u * bmK - w * bnK};
        const double U[8] = {r, r * u, r * v, r * w, e, bnK, btK, bmK};
        const double Us[8] = {rs, rs * SM, rs * vs, rs * ws, es, bn, bts, bms};
        const double Uss[8] = {rs, rs * SM, rs * vLLL, rs * wLLL, ess, bn, bttt, bmmm};
        if (outer) {
#pragma unroll
            for (int i = 0; i < 8; ++i) P.f[i] = F[i];
        } else if (single) {
#pragma unroll
            for (int i = 0; i < 8; ++i) P.f[i] = F[i] + S1 * Us[i] - S1 * U[i];
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) P.f[i] = F[i] + S2 * Uss[i] - (S2 - S1) * Us[i] - S1 * U[i];
        }
    }
    P.f[5] = -smax * (bn2 - bn1);
    return P;
}

struct MhdArgs {
    const double* Uin;
    const double* Ubase;
    double* Uout;
    const double2* s;
    const double2* u;
    const double3* b;
    double2* s2;
    double2* u2;
    double3* b2;
    int nx, ny;
    long long plane;
    double dx, dy, idx, idy;
    double gamma, g1, inv_g1, krit;
    int bc[4];
    double a, b_coef;
    int plain;
    double dt;
    const double* dt_dev;
    unsigned long long* wmax_bits;
    unsigned long long* tmin_bits;
    DeviceDiag* diag;
    int tiles_x, rows;
    // production options of the reference-layout steps (kernel.cu:1918,
    // 2137-2190): inner frozen sphere, axisymmetric sources, INFLOW ghosts
    // with B = 0 (Kernel_TVD) instead of the edge cell's (add2)
    double x0, y0, r_frozen, inv_gamma;
    int axisym, inflow_b_zero;
};

// Primitives of cell k of a SoA conserved MHD state (P = plane size).
__device__ __forceinline__ PrimM cons_to_prim_m(const double* U, long long P, long long k, double g1) {
    PrimM q;
    const double r = U[k], mx = U[P + k], my = U[2 * P + k];
    q.bx = U[3 * P + k];
    q.by = U[4 * P + k];
    q.bz = U[5 * P + k];
    const double E = U[6 * P + k];
    const double inv = fast_rcp(r);
    q.r = r;
    q.u = mx * inv;
    q.v = my * inv;
    q.p = g1 * (E - 0.5 * (mx * q.u + my * q.v) - (q.bx * q.bx + q.by * q.by + q.bz * q.bz) * (1.0 / kCpi8));
    return q;
}

template <int LAYOUT>
__device__ __forceinline__ PrimM load_cell_m(const MhdArgs& A, long long k, bool fx, bool fy) {
    PrimM q;
    if constexpr (LAYOUT == kSoA) {
        q = cons_to_prim_m(A.Uin, A.plane, k, A.g1);
    } else {
        const double2 sv = A.s[k], uv = A.u[k];
        const double3 bv = A.b[k];
        q.r = sv.x; q.p = sv.y; q.u = uv.x; q.v = uv.y; q.bx = bv.x; q.by = bv.y; q.bz = bv.z;
    }
    // reflective sides negate the normal velocity and the normal field
    if (fx) { q.u = -q.u; q.bx = -q.bx; }
    if (fy) { q.v = -q.v; q.by = -q.by; }
    return q;
}

// Logical index k -> stored index for the production side kinds of the
// reference-layout steps; side = the kind of the side k lies beyond (-1
// inside). SYMMETRY / AXIS / REFLECTIVE mirror, the others clamp.
__device__ __forceinline__ int map_index_side(int k, int n, int lo, int hi, int& side) {
    side = -1;
    if (k < 0) {
        if (lo == GD2_BC_PERIODIC) return ((k % n) + n) % n;
        side = lo;
        if (lo == GD2_BC_REFLECTIVE || lo == GD2_BC_SYMMETRY || lo == GD2_BC_AXIS) {
            const int m = -1 - k;
            return m < n ? m : n - 1;
        }
        return 0;
    }
    if (k >= n) {
        if (hi == GD2_BC_PERIODIC) return k % n;
        side = hi;
        if (hi == GD2_BC_REFLECTIVE || hi == GD2_BC_SYMMETRY || hi == GD2_BC_AXIS) {
            const int m = 2 * n - 1 - k;
            return m >= 0 ? m : 0;
        }
        return n - 1;
    }
    return k;
}

// Ghost state of a cell reached through side kinds sx (x) and sy (y), the
// rules of add2 (kernel.cu:1938-2016) and Kernel_TVD (kernel.cu:2227-2405).
__device__ __forceinline__ PrimM ghost_m(PrimM q, int sx, int sy, const MhdArgs& A) {
    if (sx == GD2_BC_REFLECTIVE) { q.u = -q.u; q.bx = -q.bx; }
    else if (sx == GD2_BC_SYMMETRY) { q.u = -q.u; q.bx = -q.bx; q.by = -q.by; }
    else if (sx == GD2_BC_OUTFLOW) { if (q.u < 0.0) q.u = 0.01; }
    if (sy == GD2_BC_REFLECTIVE) { q.v = -q.v; q.by = -q.by; }
    else if (sy == GD2_BC_AXIS) { q.v = -q.v; q.bz = -q.bz; }
    else if (sy == GD2_BC_INFLOW) {
        q.r = 1.0;
        q.p = A.inv_gamma;
        q.u = 0.0;
        q.v = 0.0;
        if (A.inflow_b_zero) { q.bx = 0.0; q.by = 0.0; q.bz = 0.0; }
    }
    return q;
}

constexpr int kMF = 7;  // LDS fields per cell: r, u, v, p, bx, by, bz

__device__ __forceinline__ void lds_put_m(double (*ring)[kMF][kRowBuf], int slot, int e, const PrimM& q) {
    ring[slot][0][e] = q.r;
    ring[slot][1][e] = q.u;
    ring[slot][2][e] = q.v;
    ring[slot][3][e] = q.p;
    ring[slot][4][e] = q.bx;
    ring[slot][5][e] = q.by;
    ring[slot][6][e] = q.bz;
}

__device__ __forceinline__ PrimM lds_get_m(double (*ring)[kMF][kRowBuf], int slot, int e) {
    PrimM q;
    q.r = ring[slot][0][e];
    q.u = ring[slot][1][e];
    q.v = ring[slot][2][e];
    q.p = ring[slot][3][e];
    q.bx = ring[slot][4][e];
    q.by = ring[slot][5][e];
    q.bz = ring[slot][6][e];
    return q;
}

template <int ORDER>
__device__ __forceinline__ PrimM slope3m(const PrimM& a, const PrimM& b, const PrimM& c) {
    PrimM s;
    if constexpr (ORDER == 2) {
        s.r = minmod(b.r - a.r, c.r - b.r);
        s.u = minmod(b.u - a.u, c.u - b.u);
        s.v = minmod(b.v - a.v, c.v - b.v);
        s.p = minmod(b.p - a.p, c.p - b.p);
        s.bx = minmod(b.bx - a.bx, c.bx - b.bx);
        s.by = minmod(b.by - a.by, c.by - b.by);
        s.bz = minmod(b.bz - a.bz, c.bz - b.bz);
    } else {
        s.r = s.u = s.v = s.p = s.bx = s.by = s.bz = 0.0;
    }
    return s;
}

__device__ __forceinline__ PrimM face_m(const PrimM& q, const PrimM& s, double h) {
    PrimM f;
    f.r = q.r + h * s.r;
    f.u = q.u + h * s.u;
    f.v = q.v + h * s.v;
    f.p = q.p + h * s.p;
    f.bx = q.bx + h * s.bx;
    f.by = q.by + h * s.by;
    f.bz = q.bz + h * s.bz;
    if (f.r <= 0.0) f.r = q.r;  // grad_p fall-back (kernel.cu:2407-2440)
    if (f.p <= 0.0) f.p = q.p;
    return f;
}

// Lab-frame face flux (7 conserved fields + normal-field flux) for an x-face
// (frame n = x, t = y, m = z) or a y-face (n = y, t = z, m = x); Gaussian field
// fluxes = sqrt(4 pi) * HL fluxes (Gas_Dinamic.cu:1285-1299).
struct LabFluxM {
    double r, mx, my, bx, by, bz, e;
};

template <int METOD, bool XFACE>
__device__ __forceinline__ LabFluxM mhd_flux(const PrimM& L, const PrimM& R, double gamma, double inv_g1,
                                             double& smax) {
    const double is = 1.0 / kSpi4;
    FluxM P;
    if constexpr (XFACE) {
        P = hlld_face<METOD>(L.r, L.p, L.u, L.v, 0.0, L.bx * is, L.by * is, L.bz * is, R.r, R.p, R.u, R.v, 0.0,
                             R.bx * is, R.by * is, R.bz * is, gamma, inv_g1, smax);
    } else {
        P = hlld_face<METOD>(L.r, L.p, L.v, 0.0, L.u, L.by * is, L.bz * is, L.bx * is, R.r, R.p, R.v, 0.0, R.u,
                             R.by * is, R.bz * is, R.bx * is, gamma, inv_g1, smax);
    }
    LabFluxM F;
    F.r = P.f[0];
    F.e = P.f[4];
    if constexpr (XFACE) {
        F.mx = P.f[1];
        F.my = P.f[2];
        F.bx = kSpi4 * P.f[5];
        F.by = kSpi4 * P.f[6];
        F.bz = kSpi4 * P.f[7];
    } else {
        F.my = P.f[1];
        F.mx = P.f[3];
        F.by = kSpi4 * P.f[5];
        F.bz = kSpi4 * P.f[6];
        F.bx = kSpi4 * P.f[7];
    }
    return F;
}

__device__ __forceinline__ LabFluxM shfl_down_m(const LabFluxM& F) {
    LabFluxM G;
    G.r = __shfl_down(F.r, 1, kWave);
    G.mx = __shfl_down(F.mx, 1, kWave);
    G.my = __shfl_down(F.my, 1, kWave);
    G.bx = __shfl_down(F.bx, 1, kWave);
    G.by = __shfl_down(F.by, 1, kWave);
    G.bz = __shfl_down(F.bz, 1, kWave);
    G.e = __shfl_down(F.e, 1, kWave);
    return G;
}

// Mirror image across an x- (y-) wall: normal velocity and normal field negated.
__device__ __forceinline__ PrimM mirror_x(PrimM q) {
    q.u = -q.u;
    q.bx = -q.bx;
    return q;
}

__device__ __forceinline__ PrimM mirror_y(PrimM q) {
    q.v = -q.v;
    q.by = -q.by;
    return q;
}

// Flux of the mirrored problem mapped back: the components odd under the
// mirror change sign (mass, tangential momentum, energy, tangential field).
__device__ __forceinline__ LabFluxM unmirror_x(LabFluxM F) {
    F.r = -F.r;
    F.my = -F.my;
    F.by = -F.by;
    F.bz = -F.bz;
    F.e = -F.e;
    return F;
}

__device__ __forceinline__ LabFluxM unmirror_y(LabFluxM F) {
    F.r = -F.r;
    F.mx = -F.mx;
    F.bx = -F.bx;
    F.bz = -F.bz;
    F.e = -F.e;
    return F;
}

template <int LAYOUT>
__device__ __forceinline__ bool frozen_m(const MhdArgs& A, int c, int j) {
    if constexpr (LAYOUT == kAoS) {
        return (A.bc[0] == GD2_BC_FROZEN && c == 0) || (A.bc[1] == GD2_BC_FROZEN && c == A.nx - 1) ||
               (A.bc[2] == GD2_BC_FROZEN && j == 0) || (A.bc[3] == GD2_BC_FROZEN && j == A.ny - 1);
    } else {
        return false;
    }
}

// Raw values of one cell as stored: the 7 conserved values (kSoA) or the
// primitives (rho, p, vx, vy, bx, by, bz) (kAoS).
struct RawM {
    double a[7];
};

// Primitives from the 7 conserved values (the expressions of cons_to_prim_m).
__device__ __forceinline__ PrimM prim_from_raw_m(const RawM& w, double g1) {
    PrimM q;
    const double r = w.a[0], mx = w.a[1], my = w.a[2];
    q.bx = w.a[3];
    q.by = w.a[4];
    q.bz = w.a[5];
    const double E = w.a[6];
    const double inv = fast_rcp(r);
    q.r = r;
    q.u = mx * inv;
    q.v = my * inv;
    q.p = g1 * (E - 0.5 * (mx * q.u + my * q.v) - (q.bx * q.bx + q.by * q.by + q.bz * q.bz) * (1.0 / kCpi8));
    return q;
}

// Loads of cell ci of stored row jj. kSoA: buffer loads from the seven
// conserved planes at byte offset `off` (kOffNone loads zeros and touches no
// memory), issued by the whole wave without a branch; kAoS: plain loads of the
// three arrays where `valid`.
template <int LAYOUT>
__device__ __forceinline__ RawM load_raw_m(const MhdArgs& A, int jj, int ci, unsigned off, bool valid) {
    RawM w;
    if constexpr (LAYOUT == kSoA) {
        const double* b = A.Uin + (long long)jj * A.nx;
#pragma unroll
        for (int f = 0; f < kMF; ++f) w.a[f] = ld_buf(buf_rsrc(uniform_ptr(b + f * A.plane)), off);
    } else {
#pragma unroll
        for (int f = 0; f < kMF; ++f) w.a[f] = 0.0;
        if (valid) {
            const long long k = (long long)jj * A.nx + ci;
            const double2 sv = A.s[k], uv = A.u[k];
            const double3 bv = A.b[k];
            w.a[0] = sv.x;
            w.a[1] = sv.y;
            w.a[2] = uv.x;
            w.a[3] = uv.y;
            w.a[4] = bv.x;
            w.a[5] = bv.y;
            w.a[6] = bv.z;
        }
    }
    return w;
}

// Primitive state of a loaded cell with its boundary rule: the flips of a
// reflective side (normal velocity and normal field negated), or with PROD
// the reference's production side kinds (ghost_m).
template <int LAYOUT, bool PROD>
__device__ __forceinline__ PrimM finish_m(const MhdArgs& A, const RawM& w, bool fx, bool fy, int sx, int sy) {
    PrimM q;
    if constexpr (LAYOUT == kSoA) {
        q = prim_from_raw_m(w, A.g1);
    } else {
        q.r = w.a[0]; q.p = w.a[1]; q.u = w.a[2]; q.v = w.a[3]; q.bx = w.a[4]; q.by = w.a[5]; q.bz = w.a[6];
    }
    if constexpr (PROD) {
        q = ghost_m(q, sx, sy, A);
    } else {
        if (fx) { q.u = -q.u; q.bx = -q.bx; }
        if (fy) { q.v = -q.v; q.by = -q.by; }
    }
    return q;
}

// Per-wave LDS of the stage kernel: primitive rows (own cells and halo) for
// the x-faces, row r in slot r & 1, and (kSoA) the own conserved values
// waiting for their update, row r in slot r % 3 (read after the barrier that
// hands the primitive slot to row r + 2): 18.4 KB.
struct MhdLds {
    double q[2][kMF][kRowBuf];
    double u[3][kMF][kWave];
};

// PROD (reference layout only): production side kinds, inner frozen sphere
// and axisymmetric sources.
//
// March of one wave (as in the Euler stage kernel): at pass j the loads of
// row j+2 are issued first and arrive while the x-face of row j is computed
// from LDS; after the barrier row j+2 is converted into the free LDS slot,
// the y-face G(j+1/2) is formed from rows j-1 .. j+2 of the lane's column,
// the RK base of row j (loaded after the barrier) arrives during it, and row
// j is updated. Every face is evaluated once; every expression is unchanged
// by the order of the march.
template <int ORDER, int METOD, int LAYOUT, bool WANT_W, bool PROD = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void mhd_tile_kernel(const MhdArgs A) {
    __shared__ MhdLds lds;

    const int lane = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % A.tiles_x, ty = tile / A.tiles_x;
    const int cbase = tx * kTileW;
    const int j0 = ty * A.rows;
    const int j1 = min(j0 + A.rows, A.ny);
    const int c = cbase + lane;

    bool fx = false, hfx = false;
    int sxc = -1, sxh = -1;
    const int hcol = lane == 0 ? cbase - 2 : (lane == 1 ? cbase - 1 : cbase + 64);
    int ci, hci;
    if constexpr (PROD) {
        ci = map_index_side(c, A.nx, A.bc[0], A.bc[1], sxc);
        hci = map_index_side(hcol, A.nx, A.bc[0], A.bc[1], sxh);
    } else {
        ci = map_index(c, A.nx, A.bc[0], A.bc[1], fx);
        hci = map_index(hcol, A.nx, A.bc[0], A.bc[1], hfx);
    }
    const int hslot = lane == 0 ? 0 : (lane == 1 ? 1 : 66);
    const bool has_halo = lane < 3;
    const bool owns = (lane < kTileW) && (c < A.nx);
    const unsigned oc = (unsigned)ci * 8u;                      // own cell within a row (kSoA)
    const unsigned oh = has_halo ? (unsigned)hci * 8u : kOffNone;  // halo cell within a row (kSoA)
    const unsigned oo = (unsigned)c * 8u;                       // output cell within a row (kSoA)
    const unsigned ob_off = owns ? oo : kOffNone;               // RK base cell within a row (kSoA)

    const double dt = A.dt_dev ? *A.dt_dev : A.dt;
    const double dx = A.dx, dy = A.dy;
    const double dV = dx * dy;
    // lagged face time step radius: dx on x-faces and dy on y-faces as in
    // add2 / Kernel_TVD (kernel.cu:2070, 2086)
    const double krx = A.krit * dx, kry = A.krit * dy;

    // logical row jl: raw loads of the own cell and (lanes 0-2) the halo cell
    struct Pending {
        RawM own, halo;
        bool fy;
        int sy;
    };
    auto issue_row = [&](int jl) -> Pending {
        Pending pr;
        pr.fy = false;
        pr.sy = -1;
        int jj;
        if constexpr (PROD)
            jj = map_index_side(jl, A.ny, A.bc[2], A.bc[3], pr.sy);
        else
            jj = map_index(jl, A.ny, A.bc[2], A.bc[3], pr.fy);
        pr.own = load_raw_m<LAYOUT>(A, jj, ci, oc, true);
        pr.halo = load_raw_m<LAYOUT>(A, jj, hci, oh, has_halo);
        return pr;
    };
    // convert a pending row; to_lds: publish it (and the halo) in LDS slot
    // `slot`, with the own conserved values in slot `uslot` (kSoA)
    auto finish_row = [&](const Pending& pr, int slot, int uslot, bool to_lds) -> PrimM {
        const PrimM q = finish_m<LAYOUT, PROD>(A, pr.own, fx, pr.fy, sxc, pr.sy);
        if (to_lds) {
            lds_put_m(lds.q, slot, lane + 2, q);
            if constexpr (LAYOUT == kSoA) {
#pragma unroll
                for (int f = 0; f < kMF; ++f) lds.u[uslot][f][lane] = pr.own.a[f];
            }
            if (has_halo) lds_put_m(lds.q, slot, hslot, finish_m<LAYOUT, PROD>(A, pr.halo, hfx, pr.fy, sxh, pr.sy));
        }
        return q;
    };

    // prologue: rows j0-2, j0-1 (y-stencil only) and j0 (also in LDS); the
    // first y-face comes from the loop's lead-in pass j = j0-1
    PrimM q0, q1, s0;
    double by_s;  // cell-centred By of the south neighbour (Powell average)
    {
        const Pending pm2 = issue_row(j0 - 2), pm1 = issue_row(j0 - 1), pj0 = issue_row(j0);
        const PrimM qm2 = finish_row(pm2, 0, 0, false);
        q0 = finish_row(pm1, 0, 0, false);
        q1 = finish_row(pj0, j0 & 1, j0 % 3, true);
        s0 = slope3m<ORDER>(qm2, q0, q1);
        by_s = qm2.by;
    }
    double smax;
    LabFluxM Gm;
    Gm.r = Gm.mx = Gm.my = Gm.bx = Gm.by = Gm.bz = Gm.e = 0.0;
    double sGm = 0.0;

    double wmax = 0.0;
    // kAoS: the largest wave speed over the x-faces and over the y-faces of
    // the cells this lane updates; the face time step krit rad / S is
    // monotone in S, so the minimum over the faces is krit rad / max S
    // (bitwise), formed once at the end
    double sx = 0.0, sy = 0.0;
    bool any_upd = false;

    // the wave behind on its SIMD takes the higher issue priority, decided
    // every kTurnPeriod passes
    SimdTurn turn;
    turn.init();
    for (int j = j0 - 1; j < j1; ++j) {
        if (((j - j0 + 1) & (kTurnPeriod - 1)) == 0) turn.step((j - j0 + 1) / kTurnPeriod);
        // loads first: row j+2 arrives while the x-face of row j is computed
        const Pending pn = issue_row(j + 2);

        // x-face F(c-1/2) from cells c-2, c-1, c, c+1 of row j (LDS; the own
        // cell is q0); F(c+1/2) from the next lane. Only the differences
        // F(c+1/2) - F(c-1/2) are carried to the update.
        LabFluxM dF;
        double sFl = 0.0, sFr = 0.0, dbx = 0.0;
        if (j >= j0) {
            const int sl = j & 1;
            const PrimM qa = lds_get_m(lds.q, sl, lane);
            const PrimM qb = lds_get_m(lds.q, sl, lane + 1);
            const PrimM qd = lds_get_m(lds.q, sl, lane + 3);
            const PrimM sb = slope3m<ORDER>(qa, qb, q0);
            const PrimM sc = slope3m<ORDER>(qb, q0, qd);
            // reflective west wall face from the interior side (see the south wall)
            const bool wmir = (c == 0) && (A.bc[0] == GD2_BC_REFLECTIVE || (PROD && A.bc[0] == GD2_BC_SYMMETRY));
            const PrimM xl = face_m(qb, sb, 0.5), xr = face_m(q0, sc, -0.5);
            LabFluxM Fl = mhd_flux<METOD, true>(wmir ? mirror_x(xr) : xl, wmir ? mirror_x(xl) : xr, A.gamma,
                                                A.inv_g1, smax);
            if (wmir) Fl = unmirror_x(Fl);
            sFl = (LAYOUT == kAoS) ? smax : 0.0;
            const LabFluxM Fr = shfl_down_m(Fl);
            sFr = (LAYOUT == kAoS) ? __shfl_down(sFl, 1, kWave) : 0.0;
            dF.r = Fr.r - Fl.r;
            dF.mx = Fr.mx - Fl.mx;
            dF.my = Fr.my - Fl.my;
            dF.bx = Fr.bx - Fl.bx;
            dF.by = Fr.by - Fl.by;
            dF.bz = Fr.bz - Fl.bz;
            dF.e = Fr.e - Fl.e;
            dbx = qd.bx - qb.bx;  // east minus west cell-centred Bx
        }

        // row j's primitive slot is read: row j+2 may take it
        __syncthreads();
        const PrimM q2 = finish_row(pn, (j + 2) & 1, (j + 2) % 3, true);
        // the RK base of row j (kSoA), in flight during the y-face; the
        // lead-in pass and lanes without a cell load nothing
        double ub[kMF];
        if constexpr (LAYOUT == kSoA) {
            if (!A.plain) {
                const unsigned off = j >= j0 ? ob_off : kOffNone;
                const double* b = A.Ubase + (long long)max(j, j0) * A.nx;
#pragma unroll
                for (int f = 0; f < kMF; ++f) ub[f] = ld_buf(buf_rsrc(uniform_ptr(b + f * A.plane)), off);
            }
        }
        const PrimM s1 = slope3m<ORDER>(q0, q1, q2);
        // A reflective south wall face is evaluated from the interior cell's
        // side (outward normal), as the reference's cell does; its HLLC-MHD /
        // HLLD flux is discontinuous at SM = 0 when the normal field jumps.
        const bool smir = (j == -1) && (A.bc[2] == GD2_BC_REFLECTIVE || (PROD && A.bc[2] == GD2_BC_AXIS));
        const PrimM yl = face_m(q0, s0, 0.5), yr = face_m(q1, s1, -0.5);
        const LabFluxM Gp = smir ? unmirror_y(mhd_flux<METOD, false>(mirror_y(yr), mirror_y(yl), A.gamma, A.inv_g1, smax))
                                 : mhd_flux<METOD, false>(yl, yr, A.gamma, A.inv_g1, smax);
        const double sGp = (LAYOUT == kAoS) ? smax : 0.0;
        if (j < j0) {  // lead-in pass: only G(j0-1/2)
            by_s = q0.by;
            q0 = q1;
            q1 = q2;
            s0 = s1;
            Gm = Gp;
            sGm = sGp;
            continue;
        }

        bool sphere = false;
        double yc = 0.0;
        if constexpr (PROD) {
            const double xc = A.x0 + c * dx;
            yc = A.y0 + j * dy;
            sphere = A.r_frozen > 0.0 && sqrt(xc * xc + yc * yc) < A.r_frozen;
        }
        if (owns && !frozen_m<LAYOUT>(A, c, j) && !sphere) {
            // face sums differenced per direction (kernel.cu:2559-2604)
            const double PSr = dF.r * dy + (Gp.r - Gm.r) * dx;
            const double PSe = dF.e * dy + (Gp.e - Gm.e) * dx;
            const double PUx = dF.mx * dy + (Gp.mx - Gm.mx) * dx;
            const double PUy = dF.my * dy + (Gp.my - Gm.my) * dx;
            const double PBx = dF.bx * dy + (Gp.bx - Gm.bx) * dx;
            const double PBy = dF.by * dy + (Gp.by - Gm.by) * dx;
            const double PBz = dF.bz * dy + (Gp.bz - Gm.bz) * dx;
            // Powell term: face-averaged normal field of the cell-centred
            // neighbours (kernel.cu:2568, 2580, 2592, 2604)
            const double Pdiv = dy * dbx * 0.5 + dx * (q1.by - by_s) * 0.5;
            const PrimM& q = q0;
            if constexpr (LAYOUT == kSoA) {
                double uc[kMF];
#pragma unroll
                for (int f = 0; f < kMF; ++f) uc[f] = lds.u[j % 3][f][lane];
                const double dtV = dt / dV;
                double o[7];
                o[0] = uc[0] - dtV * PSr;
                o[1] = uc[1] - dtV * (PUx + (q.bx * (1.0 / kCpi4)) * Pdiv);
                o[2] = uc[2] - dtV * (PUy + (q.by * (1.0 / kCpi4)) * Pdiv);
                o[3] = uc[3] - dtV * (PBx + q.u * Pdiv);
                o[4] = uc[4] - dtV * (PBy + q.v * Pdiv);
                o[5] = uc[5] - dtV * PBz;
                o[6] = uc[6] - dtV * (PSe + ((q.u * q.bx + q.v * q.by) * (1.0 / kCpi4)) * Pdiv);
                if (!A.plain) {
#pragma unroll
                    for (int f = 0; f < 7; ++f) o[f] = A.a * ub[f] + A.b_coef * o[f];
                }
                // clamps (kernel.cu:2144-2148 rho -> old rho; 2192-2195 p -> 1e-6)
                if (o[0] <= 0.0) {
                    atomicAdd(&A.diag->clamp_rho, 1ull);
                    o[0] = uc[0];
                }
                const double inv = fast_rcp(o[0]);
                const double ux = o[1] * inv, uy = o[2] * inv;
                const double ke = 0.5 * (o[1] * ux + o[2] * uy);
                const double b2 = o[3] * o[3] + o[4] * o[4] + o[5] * o[5];
                const double me = b2 * (1.0 / kCpi8);
                double p = A.g1 * (o[6] - ke - me);
                if (p <= 0.0) {
                    atomicAdd(&A.diag->clamp_p, 1ull);
                    o[6] = 0.000001 * A.inv_g1 + ke + me;
                    p = A.g1 * (o[6] - ke - me);
                }
                double* ob = A.Uout + (long long)j * A.nx;
#pragma unroll
                for (int f = 0; f < 7; ++f) st_buf(buf_rsrc(uniform_ptr(ob + f * A.plane)), oo, o[f]);
                if constexpr (WANT_W) {
                    const double cf = fast_sqrt((A.gamma * p + b2 * (1.0 / kCpi4)) * inv);
                    wmax = fmax(wmax, (fabs(ux) + cf) * A.idx + (fabs(uy) + cf) * A.idy);
                }
            } else {
                const long long k = (long long)j * A.nx + c;
                // primitive-form update of add2 / Kernel_TVD (kernel.cu:2137-2195);
                // the 1/y terms only with the axisymmetric option. Divisions by
                // dV, y, the new density and the constants as products with
                // reciprocals (dt/dV once per launch, dt/y once per row, 1/rho'
                // once per cell): rounding-level against the reference's
                // quotients.
                const double dtV = dt / dV;
                const double U8 = q.p * A.inv_g1 + 0.5 * q.r * (q.u * q.u + q.v * q.v) +
                                  (q.bx * q.bx + q.by * q.by + q.bz * q.bz) * (1.0 / kCpi8);
                const double skk = q.u * q.bx + q.v * q.by;
                double r2, ux, uy, bx, by, bz, p2;
                if (PROD && A.axisym) {
                    const double iy = 1.0 / yc, dty = dt * iy;
                    const double Pd = Pdiv + dV * q.by * iy;
                    r2 = q.r - dtV * PSr - dty * q.r * q.v;
                    if (r2 <= 0.0) {
                        atomicAdd(&A.diag->clamp_rho, 1ull);
                        r2 = q.r;
                    }
                    const double ir2 = 1.0 / r2;
                    ux = (q.r * q.u - dtV * (PUx + (q.bx * (1.0 / kCpi4)) * Pd) -
                          dty * (q.r * q.u * q.v - q.bx * q.by * (1.0 / kCpi4))) * ir2;
                    uy = (q.r * q.v - dtV * (PUy + (q.by * (1.0 / kCpi4)) * Pd) -
                          dty * (q.r * q.v * q.v + (q.bz * q.bz - q.by * q.by) * (1.0 / kCpi4))) * ir2;
                    bx = q.bx - dtV * (PBx + q.u * Pd) - dty * (q.v * q.bx - q.by * q.u);
                    by = q.by - dtV * (PBy + q.v * Pd);
                    bz = q.bz - dtV * PBz;
                    p2 = (U8 - dtV * (PSe + (skk * (1.0 / kCpi4)) * Pd) -
                          dty * ((U8 + q.p + (q.bx * q.bx + q.by * q.by + q.bz * q.bz) * (1.0 / kCpi8)) * q.v -
                                 q.by * skk * (1.0 / kCpi4)) -
                          0.5 * r2 * (ux * ux + uy * uy) - (bx * bx + by * by + bz * bz) * (1.0 / kCpi8)) *
                         (A.gamma - 1.0);
                } else {
                    r2 = q.r - dtV * PSr;
                    if (r2 <= 0.0) {
                        atomicAdd(&A.diag->clamp_rho, 1ull);
                        r2 = q.r;
                    }
                    const double ir2 = 1.0 / r2;
                    ux = (q.r * q.u - dtV * (PUx + (q.bx * (1.0 / kCpi4)) * Pdiv)) * ir2;
                    uy = (q.r * q.v - dtV * (PUy + (q.by * (1.0 / kCpi4)) * Pdiv)) * ir2;
                    bx = q.bx - dtV * (PBx + q.u * Pdiv);
                    by = q.by - dtV * (PBy + q.v * Pdiv);
                    bz = q.bz - dtV * PBz;
                    p2 = (U8 - dtV * (PSe + (skk * (1.0 / kCpi4)) * Pdiv) - 0.5 * r2 * (ux * ux + uy * uy) -
                          (bx * bx + by * by + bz * bz) * (1.0 / kCpi8)) * (A.gamma - 1.0);
                }
                if (p2 <= 0.0) {
                    atomicAdd(&A.diag->clamp_p, 1ull);
                    p2 = 0.000001;
                }
                A.s2[k] = make_double2(r2, p2);
                A.u2[k] = make_double2(ux, uy);
                A.b2[k] = make_double3(bx, by, bz);
                sx = fmax(sx, fmax(sFr, sFl));
                sy = fmax(sy, fmax(sGm, sGp));
                any_upd = true;
            }
        } else if constexpr (LAYOUT == kAoS) {
            if (owns) {
                const long long k = (long long)j * A.nx + c;
                A.s2[k] = A.s[k];
                A.u2[k] = A.u[k];
                A.b2[k] = A.b[k];
            }
        }

        by_s = q0.by;
        q0 = q1;
        q1 = q2;
        s0 = s1;
        Gm = Gp;
        sGm = sGp;
    }

    turn.finish();
    if constexpr (LAYOUT == kSoA && WANT_W) {
        const double w = wave_max(wmax);
        if (lane == 0 && A.wmax_bits) atomic_max_nonneg(A.wmax_bits, w);
    }
    if constexpr (LAYOUT == kAoS) {
        const double tmin = any_upd ? fmin(1000.0, fmin(krx / sx, kry / sy)) : 1e300;
        const double t = wave_min(tmin);
        if (lane == 0 && t < 1e300) atomic_min_nonneg(A.tmin_bits, t);
    }
}

static __global__ __launch_bounds__(256) void mhd_wavespeed_kernel(const double* U, long long n, double gamma, double g1,
                                                            double idx, double idy, unsigned long long* wmax_bits) {
    double w = 0.0;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        const double r = U[k], mx = U[n + k], my = U[2 * n + k];
        const double bx = U[3 * n + k], by = U[4 * n + k], bz = U[5 * n + k], E = U[6 * n + k];
        const double inv = fast_rcp(r);
        const double ux = mx * inv, uy = my * inv;
        const double b2 = bx * bx + by * by + bz * bz;
        const double p = g1 * (E - 0.5 * (mx * ux + my * uy) - b2 * (1.0 / kCpi8));
        const double cf = fast_sqrt((gamma * p + b2 * (1.0 / kCpi4)) * inv);
        w = fmax(w, (fabs(ux) + cf) * idx + (fabs(uy) + cf) * idy);
    }
    __shared__ double part[256 / kWave];
    w = wave_max(w);
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = part[0];
        for (int i = 1; i < (int)(blockDim.x / kWave); ++i) m = fmax(m, part[i]);
        atomic_max_nonneg(wmax_bits, m);
    }
}

}  // namespace gd2
