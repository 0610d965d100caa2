// gd2_gkr.hpp -- the Godunov-Kolgan-Rodionov predictor-corrector MHD step
// (Kernel_GKR, kernel.cu:2663-3284, with predictor kernel.cu:1611-1720 and
// POTOK_Korolkov Gas_Dinamic.cu:70-181) in the reference AoS layout.
//
// The reference evaluates the step one thread per cell: the cell gathers its
// 13-point stencil (radius 2 along the axes plus the four diagonals) through
// the step's own side rules, forms minmod slopes of itself and its four
// neighbours, a half-step predictor of each of the five (physical fluxes of
// the half-slope face states plus the axisymmetric sources), and updates
// itself from the mean of its state and its predictor with the HLLD (or
// HLLC-MHD) fluxes of the predictor-averaged face states, over dt/2 with the
// axisymmetric sources; each face from the cell's own side with its outward
// normal. gkr_cell_step restates that for the two-cell band next to the
// sides, where the side rules make a neighbour's values depend on the cell
// that reads them; the interior, where no side rule applies, is a march
// (gkr_march_kernel) that forms each cell's slopes and predictor once and
// each face's flux once. Divisions are products with reciprocals.
//
// State vector Q7 = (rho, p, vx, vy, Bx, By, Bz), Gaussian B.
#pragma once

#include "gd2_device.hpp"
#include "gd2_mhd.hpp"

namespace gd2 {

struct GkrArgs {
    const double2* s;
    const double2* u;
    const double3* b;
    double2* s2;
    double2* u2;
    double3* b2;
    unsigned long long* tmin_bits;  // the reference's *T, folded as the bits of a non-negative double
    const double* T_do;
    int nx, ny;
    double dx, dy, x0, y0, gamma, inv_g1, krit, r_frozen;
    double idV;      // 1 / (dx dy)
    int tiles_x, rows, march_blocks;  // strips of the interior march, then the band blocks
    DeviceDiag* diag;
};

struct Q7 {
    double v[7];
};

__device__ __forceinline__ Q7 gkr_cell(const GkrArgs& A, int n, int m) {
    const long long k = (long long)m * A.nx + n;
    const double2 s = A.s[k], u = A.u[k];
    const double3 b = A.b[k];
    Q7 q;
    q.v[0] = s.x;
    q.v[1] = s.y;
    q.v[2] = u.x;
    q.v[3] = u.y;
    q.v[4] = b.x;
    q.v[5] = b.y;
    q.v[6] = b.z;
    return q;
}

// mirror in x (vx, Bx, By negated), in y (vy, Bz negated), the outflow rule
// (vx raised to 0.01 where negative) and the inflow state with the field of
// `bsrc` (kernel.cu:2693-2760)
__device__ __forceinline__ Q7 gkr_mx(Q7 q) {
    q.v[2] = -q.v[2];
    q.v[4] = -q.v[4];
    q.v[5] = -q.v[5];
    return q;
}
__device__ __forceinline__ Q7 gkr_my(Q7 q) {
    q.v[3] = -q.v[3];
    q.v[6] = -q.v[6];
    return q;
}
__device__ __forceinline__ Q7 gkr_out(Q7 q) {
    if (q.v[2] < 0) q.v[2] = 0.01;
    return q;
}
__device__ __forceinline__ Q7 gkr_in(double gamma, const Q7& bsrc) {
    Q7 q;
    q.v[0] = 1.0;
    q.v[1] = 1.0 / gamma;
    q.v[2] = 0.0;
    q.v[3] = 0.0;
    q.v[4] = bsrc.v[4];
    q.v[5] = bsrc.v[5];
    q.v[6] = bsrc.v[6];
    return q;
}

// The stencil points s_1 .. s_13 of kernel.cu:2678-2957 (the reference's
// numbering: 2 east, 3 south, 4 west, 5 north, 6 east-east, 7 south-south,
// 8 west-west, 9 north-north, 10 north-east, 11 south-east, 12 south-west,
// 13 north-west), with its side rules and its second-neighbour choices.
struct GkrStencil {
    const GkrArgs& A;
    int n, m;

    __device__ Q7 p1() const { return gkr_cell(A, n, m); }
    __device__ Q7 p5() const { return m == A.ny - 1 ? gkr_in(A.gamma, p1()) : gkr_cell(A, n, m + 1); }
    __device__ Q7 p2() const { return n == A.nx - 1 ? gkr_out(p1()) : gkr_cell(A, n + 1, m); }
    __device__ Q7 p4() const { return n == 0 ? gkr_mx(p1()) : gkr_cell(A, n - 1, m); }
    __device__ Q7 p3() const { return m == 0 ? gkr_my(p1()) : gkr_cell(A, n, m - 1); }
    __device__ Q7 p10() const {
        if (n == A.nx - 1) return p5();
        if (m == A.ny - 1) return gkr_in(A.gamma, p2());
        return gkr_cell(A, n + 1, m + 1);
    }
    __device__ Q7 p12() const {
        if (m == 0) return gkr_my(p4());
        if (n == 0) return gkr_mx(p3());
        return gkr_cell(A, n - 1, m - 1);
    }
    __device__ Q7 p11() const {
        if (m == 0) return gkr_my(p2());
        if (n == A.nx - 1) return gkr_out(p3());
        return gkr_cell(A, n + 1, m - 1);
    }
    __device__ Q7 p13() const {
        if (n == 0) return gkr_mx(p5());
        if (m == A.ny - 1) return gkr_in(A.gamma, p4());
        return gkr_cell(A, n - 1, m + 1);
    }
    __device__ Q7 p6() const { return n >= A.nx - 2 ? gkr_out(p2()) : gkr_cell(A, n + 2, m); }
    __device__ Q7 p9() const { return m >= A.ny - 2 ? p5() : gkr_cell(A, n, m + 2); }
    __device__ Q7 p8() const { return (n == 0 || n == 1) ? gkr_mx(p2()) : gkr_cell(A, n - 2, m); }
    __device__ Q7 p7() const {
        if (m == 1) return gkr_my(p5());
        if (m == 0) return gkr_my(p3());
        return gkr_cell(A, n, m - 2);
    }
};

// h * minmod((hi - mid)/h, (mid - lo)/h) per component (kernel.cu:2965-3038),
// formed as minmod(hi - mid, mid - lo): the same choice for h > 0, without
// the two divisions and the product (rounding-level); the vy slot takes the
// pressure slope, as the reference writes it
__device__ __forceinline__ Q7 gkr_slope(const Q7& hi, const Q7& mid, const Q7& lo, double) {
    Q7 d;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const int c = k == 3 ? 1 : k;
        d.v[k] = minmod(hi.v[c] - mid.v[c], mid.v[c] - lo.v[c]);
    }
    return d;
}

// POTOK_Korolkov (Gas_Dinamic.cu:70-181) with vz = 0 and an in-plane axis
// normal (n1, n2): the physical MHD flux in the lab frame, P = (F_rho, F_mx,
// F_my, F_mz, F_Bx, F_By, F_Bz, F_E).
__device__ __forceinline__ void gkr_potok(double inv_g1, double ro, double p, double vx, double vy, double Bx,
                                          double By, double Bz, double n1, double n2, double (&P)[8]) {
    const double bx = Bx * (1.0 / kSpi4), by = By * (1.0 / kSpi4), bz = Bz * (1.0 / kSpi4);
    double t1 = 0, t2 = 0, t3 = 0, m1 = 0, m2 = 0, m3 = 0;
    if (n1 > 0.1) {
        t2 = 1.0;
        m3 = 1.0;
    } else if (n2 > 0.1) {
        t3 = 1.0;
        m1 = 1.0;
    } else if (n1 < -0.1) {
        t3 = -1.0;
        m2 = -1.0;
    } else {
        t1 = -1.0;
        m3 = -1.0;
    }
    const double u1 = vx * n1 + vy * n2;
    const double v1 = vx * t1 + vy * t2;
    const double w1 = vx * m1 + vy * m2;
    const double bn = bx * n1 + by * n2;
    const double bt = bx * t1 + by * t2 + bz * t3;
    const double bm = bx * m1 + by * m2 + bz * m3;
    const double bb = bx * bx + by * by + bz * bz;
    const double uu = (vx * vx + vy * vy) / 2.0;
    const double pT = p + bb / 2.0;
    const double e = p * inv_g1 + ro * uu + bb / 2.0;
    const double f0 = ro * u1;
    const double f1 = ro * u1 * u1 + pT - bn * bn;
    const double f2 = ro * u1 * v1 - bn * bt;
    const double f3 = ro * u1 * w1 - bn * bm;
    const double f4 = (e + pT) * u1 - bn * (u1 * bn + v1 * bt + w1 * bm);
    const double f6 = u1 * bt - v1 * bn;
    const double f7 = u1 * bm - w1 * bn;
    P[0] = f0;
    P[1] = n1 * f1 + t1 * f2 + m1 * f3;
    P[2] = n2 * f1 + t2 * f2 + m2 * f3;
    P[3] = t3 * f2 + m3 * f3;
    P[4] = kSpi4 * (t1 * f6 + m1 * f7);
    P[5] = kSpi4 * (t2 * f6 + m2 * f7);
    P[6] = kSpi4 * (t3 * f6 + m3 * f7);
    P[7] = f4;
}

__device__ __forceinline__ double gkr_u8(double inv_g1, double ro, double p, double u, double v, double bx, double by,
                                         double bz) {
    // Header.h:16 with w = 0
    return p * inv_g1 + 0.5 * ro * (u * u + v * v) + (bx * bx + by * by + bz * bz) * (1.0 / kCpi8);
}

// The seven flux sums of a cell: F_rho, F_E, F_mx, F_my, F_Bx, F_By, F_Bz
// times the face lengths, summed over its faces.
struct GkrSums {
    double PSx, PSy, PUx, PUy, PBx, PBy, PBz;
};

// predictor (kernel.cu:1611-1720), its flux part: the physical fluxes of the
// four half-slope face states (rho, p falling back to the cell value when not
// positive), independent of the radius
__device__ inline GkrSums gkr_pred_flux(const GkrArgs& A, const Q7& Q, const Q7& DX, const Q7& DY) {
    const double dx = A.dx, dy = A.dy;
    GkrSums S = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        // west (-DX/2, -x), east (+DX/2, +x), north (+DY/2, +y), south (-DY/2, -y)
        const Q7& D = f < 2 ? DX : DY;
        const bool minus = f == 0 || f == 3;
        const double n1 = f == 0 ? -1.0 : (f == 1 ? 1.0 : 0.0);
        const double n2 = f == 2 ? 1.0 : (f == 3 ? -1.0 : 0.0);
        const double len = f < 2 ? dy : dx;
        double q[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) q[k] = minus ? Q.v[k] - D.v[k] / 2.0 : Q.v[k] + D.v[k] / 2.0;
        if (q[0] <= 0.0) q[0] = Q.v[0];
        if (q[1] <= 0.0) q[1] = Q.v[1];
        double P[8];
        gkr_potok(A.inv_g1, q[0], q[1], q[2], q[3], q[4], q[5], q[6], n1, n2, P);
        S.PSx = S.PSx + P[0] * len;
        S.PSy = S.PSy + P[7] * len;
        S.PUx = S.PUx + P[1] * len;
        S.PUy = S.PUy + P[2] * len;
        S.PBx = S.PBx + P[4] * len;
        S.PBy = S.PBy + P[5] * len;
        S.PBz = S.PBz + P[6] * len;
    }
    return S;
}

// The conserved update of kernel.cu:1690-1720 (predictor, st = dt/2 of the
// state Q) and kernel.cu:3247-3283 (corrector, from the mean state): a step of
// st from the flux sums with the axisymmetric source terms at radius y and
// Powell's term zero; the density and pressure clamps set bits 0 and 1 of
// `clamps`. Divisions by dV, y and rho as products with their reciprocals.
__device__ inline Q7 gkr_update(const GkrArgs& A, const double (&q)[7], const GkrSums& S, double st, double y,
                                unsigned& clamps) {
    const double sV = st * A.idV, sy = st * fast_rcp(y);
    Q7 QQ;
    double r = q[0] - sV * S.PSx - sy * q[0] * q[3];
    if (r <= 0) {
        clamps |= 1u;
        r = q[0];
    }
    const double ir = fast_rcp(r);
    QQ.v[0] = r;
    QQ.v[2] = (q[0] * q[2] - sV * (S.PUx) - sy * (q[0] * q[2] * q[3] - q[4] * q[5] * (1.0 / kCpi4))) * ir;
    QQ.v[3] = (q[0] * q[3] - sV * (S.PUy) - sy * (q[0] * q[3] * q[3] + (q[6] * q[6] - q[5] * q[5]) * (1.0 / kCpi4))) * ir;
    QQ.v[4] = (q[4] - sV * (S.PBx) - sy * (q[3] * q[4] - q[5] * q[2]));
    QQ.v[5] = (q[5] - sV * (S.PBy));
    QQ.v[6] = (q[6] - sV * (S.PBz));
    const double u8 = gkr_u8(A.inv_g1, q[0], q[1], q[2], q[3], q[4], q[5], q[6]);
    double p = (u8 - sV * (S.PSy) -
                sy * (((u8 + q[1] + (q[4] * q[4] + q[5] * q[5] + q[6] * q[6]) * (1.0 / kCpi8)) * q[3] -
                       q[5] * (q[2] * q[4] + q[3] * q[5]) * (1.0 / kCpi4))) -
                0.5 * r * (QQ.v[2] * QQ.v[2] + QQ.v[3] * QQ.v[3]) -
                (QQ.v[4] * QQ.v[4] + QQ.v[5] * QQ.v[5] + QQ.v[6] * QQ.v[6]) * (1.0 / kCpi8)) * (A.gamma - 1.0);
    if (p <= 0) {
        clamps |= 2u;
        p = 0.000001;
    }
    QQ.v[1] = p;
    return QQ;
}

__device__ __forceinline__ void gkr_count_clamps(const GkrArgs& A, unsigned clamps, unsigned long long users) {
    if (!A.diag || users == 0) return;
    if (clamps & 1u) atomicAdd(&A.diag->clamp_rho, users);
    if (clamps & 2u) atomicAdd(&A.diag->clamp_p, users);
}

// predictor (kernel.cu:1611-1720): a step of st from the physical fluxes of
// the four half-slope face states and the axisymmetric source terms at radius y
__device__ inline Q7 gkr_predictor(const GkrArgs& A, const Q7& Q, const Q7& DX, const Q7& DY, double st, double y) {
    unsigned clamps = 0;
    const Q7 QQ = gkr_update(A, Q.v, gkr_pred_flux(A, Q, DX, DY), st, y, clamps);
    gkr_count_clamps(A, clamps, 1);
    return QQ;
}

// HLLDQ_Korolkov (metod 2 / 3) of lab states L, R (vz = 0) through a face
// with in-plane axis normal (n1, n2), in the frame of its normal table
// (Gas_Dinamic.cu:978-1019); lab-frame flux P as in the reference after its
// SWAP (energy last), face time step krit * rad / S.
template <int METOD>
__device__ __forceinline__ double gkr_face(const GkrArgs& A, const double (&L)[7], const double (&R)[7], double n1,
                                           double n2, double rad, double (&P)[8]) {
    double t1 = 0, t2 = 0, t3 = 0, m1 = 0, m2 = 0, m3 = 0;
    if (n1 > 0.1) {
        t2 = 1.0;
        m3 = 1.0;
    } else if (n2 > 0.1) {
        t3 = 1.0;
        m1 = 1.0;
    } else if (n1 < -0.1) {
        t3 = -1.0;
        m2 = -1.0;
    } else {
        t1 = -1.0;
        m3 = -1.0;
    }
    const double bxL = L[4] * (1.0 / kSpi4), byL = L[5] * (1.0 / kSpi4), bzL = L[6] * (1.0 / kSpi4);
    const double bxR = R[4] * (1.0 / kSpi4), byR = R[5] * (1.0 / kSpi4), bzR = R[6] * (1.0 / kSpi4);
    double smax;
    const FluxM PO = hlld_face<METOD>(
        L[0], L[1], L[2] * n1 + L[3] * n2, L[2] * t1 + L[3] * t2, L[2] * m1 + L[3] * m2, bxL * n1 + byL * n2,
        bxL * t1 + byL * t2 + bzL * t3, bxL * m1 + byL * m2 + bzL * m3, R[0], R[1], R[2] * n1 + R[3] * n2,
        R[2] * t1 + R[3] * t2, R[2] * m1 + R[3] * m2, bxR * n1 + byR * n2, bxR * t1 + byR * t2 + bzR * t3,
        bxR * m1 + byR * m2 + bzR * m3, A.gamma, A.inv_g1, smax);
    P[0] = PO.f[0];
    P[1] = n1 * PO.f[1] + t1 * PO.f[2] + m1 * PO.f[3];
    P[2] = n2 * PO.f[1] + t2 * PO.f[2] + m2 * PO.f[3];
    P[3] = t3 * PO.f[2] + m3 * PO.f[3];
    P[4] = kSpi4 * (n1 * PO.f[5] + t1 * PO.f[6] + m1 * PO.f[7]);
    P[5] = kSpi4 * (n2 * PO.f[5] + t2 * PO.f[6] + m2 * PO.f[7]);
    P[6] = kSpi4 * (t3 * PO.f[6] + m3 * PO.f[7]);
    P[7] = PO.f[4];
    return A.krit * rad / smax;
}

// One cell of the GKR step as the reference evaluates it (kernel.cu:2663-3284):
// its 13-point stencil through the side rules, the slopes and predictors of
// the cell and its four neighbours, its four faces from its own side.
template <int METOD>
__device__ inline void gkr_cell_step(const GkrArgs& A, int n, int m, double& tmin) {
    const long long idx = (long long)m * A.nx + n;
    const double x = A.x0 + n * A.dx, y = A.y0 + m * A.dy;
    if (sqrt(x * x + y * y) < A.r_frozen) {
        // kernel.cu:2688-2695: frozen cells keep their state
        A.s2[idx] = A.s[idx];
        A.u2[idx] = A.u[idx];
        A.b2[idx] = A.b[idx];
    } else {
        const GkrStencil S{A, n, m};
        const double dx = A.dx, dy = A.dy;
        const double h = *A.T_do / 2.0;
        Q7 DX1, DY1;
        double Qb[7];
        {
            const Q7 q1 = S.p1(), q2 = S.p2(), q3 = S.p3(), q4 = S.p4(), q5 = S.p5();
            DX1 = gkr_slope(q2, q1, q4, dx);
            DY1 = gkr_slope(q5, q1, q3, dy);
            const Q7 QQ1 = gkr_predictor(A, q1, DX1, DY1, h, y);
#pragma unroll
            for (int k = 0; k < 7; ++k) Qb[k] = (q1.v[k] + QQ1.v[k]) / 2.0;
        }
        double PSx = 0, PSy = 0, PUx = 0, PUy = 0, PBx = 0, PBy = 0, PBz = 0;
        // east, south, west, north faces (kernel.cu:3105-3238); the
        // neighbour's slopes and predictor, its half-slope face state.
        // The stencil points are read again in each face (cache hits) from
        // indices passed through an empty asm, so that no point stays in
        // registers across the faces.
#pragma unroll 1
        for (int f = 0; f < 4; ++f) {
            int nn = n, mm = m;
            asm volatile("" : "+v"(nn), "+v"(mm));
            const GkrStencil T{A, nn, mm};
            Q7 qk, DXk, DYk;
            double yk = y;
            if (f == 0) {
                qk = T.p2();
                DXk = gkr_slope(T.p6(), qk, T.p1(), dx);
                DYk = gkr_slope(T.p10(), qk, T.p11(), dy);
            } else if (f == 1) {
                qk = T.p3();
                DXk = gkr_slope(T.p11(), qk, T.p12(), dx);
                DYk = gkr_slope(T.p1(), qk, T.p7(), dy);
                yk = y - dy;
            } else if (f == 2) {
                qk = T.p4();
                DXk = gkr_slope(T.p1(), qk, T.p8(), dx);
                DYk = gkr_slope(T.p13(), qk, T.p12(), dy);
            } else {
                qk = T.p5();
                DXk = gkr_slope(T.p10(), qk, T.p13(), dx);
                DYk = gkr_slope(T.p9(), qk, T.p1(), dy);
                yk = y + dy;
            }
            const Q7 QQk = gkr_predictor(A, qk, DXk, DYk, h, yk);
            const Q7& Din = (f == 0 || f == 2) ? DX1 : DY1;
            const Q7& Dnb = (f == 0 || f == 2) ? DXk : DYk;
            const bool plus = f == 0 || f == 3;  // the cell's side of the face is at +D/2
            double L[7], R[7];
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const double mk = (qk.v[k] + QQk.v[k]) / 2.0;
                L[k] = plus ? Qb[k] + Din.v[k] / 2.0 : Qb[k] - Din.v[k] / 2.0;
                R[k] = plus ? mk - Dnb.v[k] / 2.0 : mk + Dnb.v[k] / 2.0;
                if (k < 2) {
                    if (L[k] <= 0.0) L[k] = Qb[k];
                    if (R[k] <= 0.0) R[k] = mk;
                }
            }
            const double n1 = f == 0 ? 1.0 : (f == 2 ? -1.0 : 0.0);
            const double n2 = f == 3 ? 1.0 : (f == 1 ? -1.0 : 0.0);
            const double len = (f == 0 || f == 2) ? dy : dx;
            double P[8];
            const double t = gkr_face<METOD>(A, L, R, n1, n2, (f == 0 || f == 2) ? dx : dy, P);
            tmin = tmin <= t ? tmin : t;  // my_min, kernel.cu:102
            PSx = PSx + P[0] * len;
            PSy = PSy + P[7] * len;
            PUx = PUx + P[1] * len;
            PUy = PUy + P[2] * len;
            PBx = PBx + P[4] * len;
            PBy = PBy + P[5] * len;
            PBz = PBz + P[6] * len;
        }
        // kernel.cu:3247-3283: from the mean of the cell and its predictor,
        // dt/2, the axisymmetric sources, Powell's term zero
        unsigned clamps = 0;
        const Q7 Q2 = gkr_update(A, Qb, GkrSums{PSx, PSy, PUx, PUy, PBx, PBy, PBz}, h, y, clamps);
        gkr_count_clamps(A, clamps, 1);
        const double r2 = Q2.v[0], p2 = Q2.v[1], vx2 = Q2.v[2], vy2 = Q2.v[3], bx2 = Q2.v[4], by2 = Q2.v[5],
                     bz2 = Q2.v[6];
        A.s2[idx] = make_double2(r2, p2);
        A.u2[idx] = make_double2(vx2, vy2);
        A.b2[idx] = make_double3(bx2, by2, bz2);
    }
}

// The two-cell boundary band in order: rows 0, 1, ny-2, ny-1, then columns 0,
// 1, nx-2, nx-1 of rows 2 .. ny-3 (nx, ny >= 5); t < gkr_band_cells(A).
__device__ __forceinline__ long long gkr_band_cells(const GkrArgs& A) { return 4LL * A.nx + 4LL * (A.ny - 4); }
__device__ __forceinline__ void gkr_band_cell(const GkrArgs& A, long long t, int& n, int& m) {
    const long long rows4 = 4LL * A.nx;
    if (t < rows4) {
        const int rs = (int)(t / A.nx);
        n = (int)(t % A.nx);
        m = rs < 2 ? rs : A.ny - 4 + rs;
    } else {
        const long long u = t - rows4;
        const int cs = (int)(u & 3);
        n = cs < 2 ? cs : A.nx - 4 + cs;
        m = 2 + (int)(u >> 2);
    }
}

// Every cell per thread (grids without an interior, nx or ny < 5).
template <int METOD>
__global__ __launch_bounds__(256) void gkr_kernel(const GkrArgs A) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double tmin = 1e300;
    if (t < (long long)A.nx * A.ny) gkr_cell_step<METOD>(A, (int)(t % A.nx), (int)(t / A.nx), tmin);
    // exact device-wide minimum of the face time steps (the reference's
    // racy *T fold, kernel.cu:3241-3245)
    const double w = wave_min(tmin);
    if ((threadIdx.x & 63) == 0 && w < 1e300) atomic_min_nonneg(A.tmin_bits, w);
}

// The GKR step's interior (cells 2 <= n <= nx-3, 2 <= m <= ny-3, where no
// side rule applies) as a march: one wave per strip of kGkrW columns and
// `rows` rows marching up, lane l holding column kGkrW tx + l (lanes 2..61
// own the strip's cells). Each cell's slopes, predictor and mean state are
// formed once, when its row enters the march (the predictor at the cell's own
// radius), and each face's HLLD flux once (x-faces by the lane west of them,
// y-faces when the row above exists), given to the cell on the other side
// negated. The reference evaluates the predictors of a cell's neighbours in
// the cell's own stencil (at radius y -/+ dy for the south / north one) and
// every face from both of its cells: against that these differ by rounding
// (the flux is mirror-covariant) except where the normal field jumps at a face
// with a contact speed of exactly 0; the band next to the sides, where the
// reference's side rules make the neighbours' values depend on the cell that
// reads them, is stepped per cell by the blocks after the march's. Clamps of a predictor count once
// for each cell of the strip that would have read it; the face time step
// enters the minimum for each face of a cell that is updated here.
constexpr int kGkrW = 60;

template <int METOD>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void gkr_march_kernel(const GkrArgs A) {
    const int lane = threadIdx.x;
    if ((int)blockIdx.x >= A.march_blocks) {
        // the blocks after the march: the boundary band, one cell per thread
        const long long t = (long long)(blockIdx.x - A.march_blocks) * kWave + lane;
        double tb = 1e300;
        if (t < gkr_band_cells(A)) {
            int n, m;
            gkr_band_cell(A, t, n, m);
            gkr_cell_step<METOD>(A, n, m, tb);
        }
        const double w = wave_min(tb);
        if (lane == 0 && w < 1e300) atomic_min_nonneg(A.tmin_bits, w);
        return;
    }
    const int tile = xcd_tile(blockIdx.x, A.march_blocks);
    const int tx = tile % A.tiles_x, ty = tile / A.tiles_x;
    const int nx = A.nx;
    const int c = tx * kGkrW + lane;
    const int cc = min(c, nx - 1);  // lanes past the grid feed no cell of the strip
    const int r0 = 2 + ty * A.rows, r1 = min(r0 + A.rows, A.ny - 2);
    const bool own_col = lane >= 2 && lane <= 61 && c <= nx - 3;
    const double dx = A.dx, dy = A.dy;
    const double h = *A.T_do / 2.0;
    const double x = A.x0 + c * dx;
    auto frozen = [&](int r) {
        const double y = A.y0 + r * dy;
        return sqrt(x * x + y * y) < A.r_frozen;
    };
    // an updated cell of this strip: owned column, owned row, not frozen
    auto live = [&](int r) { return own_col && r >= r0 && r < r1 && !frozen(r); };
    auto lane_next = [](bool b) { return from_next_lane(b ? 1.0 : 0.0) != 0.0; };
    auto lane_prev = [](bool b) { return from_prev_lane(b ? 1.0 : 0.0) != 0.0; };
    auto load = [&](int r) { return gkr_cell(A, cc, r); };

    // The row data of row r at this lane: mean state m = (q + predictor) / 2
    // and its four half-slope side states (rho, p falling back to m)
    struct RowD {
        double m[7], E[7], W[7], N[7], S[7];
    };
    auto row_data = [&](const Q7& qs, const Q7& q, const Q7& qn, int r, RowD& R) -> unsigned {
        Q7 qw, qe;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            qw.v[k] = from_prev_lane(q.v[k]);
            qe.v[k] = from_next_lane(q.v[k]);
        }
        const Q7 DX = gkr_slope(qe, q, qw, dx), DY = gkr_slope(qn, q, qs, dy);
        unsigned clamps = 0;
        const Q7 QQ = gkr_update(A, q.v, gkr_pred_flux(A, q, DX, DY), h, A.y0 + r * dy, clamps);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const double mk = (q.v[k] + QQ.v[k]) / 2.0;
            R.m[k] = mk;
            R.E[k] = mk + DX.v[k] / 2.0;
            R.W[k] = mk - DX.v[k] / 2.0;
            R.N[k] = mk + DY.v[k] / 2.0;
            R.S[k] = mk - DY.v[k] / 2.0;
            if (k < 2) {
                if (R.E[k] <= 0.0) R.E[k] = mk;
                if (R.W[k] <= 0.0) R.W[k] = mk;
                if (R.N[k] <= 0.0) R.N[k] = mk;
                if (R.S[k] <= 0.0) R.S[k] = mk;
            }
        }
        return clamps;
    };
    // clamps of the predictor of row r at this lane, counted for each live
    // reader in the strip (itself, its east, west, north and south neighbours)
    auto count_pred = [&](unsigned clamps, int r) {
        const bool self = live(r);
        const bool east = lane_next(self), west = lane_prev(self);
        if (clamps) {
            const unsigned long long users =
                (self ? 1ull : 0ull) + (east ? 1ull : 0ull) + (west ? 1ull : 0ull) + (live(r + 1) ? 1ull : 0ull) +
                (live(r - 1) ? 1ull : 0ull);
            gkr_count_clamps(A, clamps, users);
        }
    };
    // the lab-frame flux sums of one face: F_rho, F_E, F_mx, F_my, F_Bx, F_By, F_Bz
    auto face = [&](const double (&L)[7], const double (&R)[7], bool xface, double (&F)[7]) -> double {
        double P[8];
        const double t = gkr_face<METOD>(A, L, R, xface ? 1.0 : 0.0, xface ? 0.0 : 1.0, xface ? dx : dy, P);
        F[0] = P[0];
        F[1] = P[7];
        F[2] = P[1];
        F[3] = P[2];
        F[4] = P[4];
        F[5] = P[5];
        F[6] = P[6];
        return t;
    };

    // LDS: the primitive rows p and p+1 (slot row & 1), and the mean state and
    // north side state of row p, which wait there across the face fluxes
    __shared__ double qrow[2][7][kWave];
    __shared__ double rowp[2][7][kWave];  // [0] m, [1] N
    auto put_q = [&](int r, const Q7& q) {
#pragma unroll
        for (int k = 0; k < 7; ++k) qrow[r & 1][k][lane] = q.v[k];
    };
    auto get_q = [&](int r, int e) {
        Q7 q;
#pragma unroll
        for (int k = 0; k < 7; ++k) q.v[k] = qrow[r & 1][k][e];
        return q;
    };

    double tmin = 1e300;
    // prologue: rows r0-2 .. r0, the row data of row r0-1
    {
        const Q7 qa = load(r0 - 2), qb = load(r0 - 1), qc = load(r0);
        put_q(r0 - 1, qb);
        put_q(r0, qc);
        RowD R;
        const unsigned cl = row_data(qa, qb, qc, r0 - 1, R);
        count_pred(cl, r0 - 1);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            rowp[0][k][lane] = R.m[k];
            rowp[1][k][lane] = R.N[k];
        }
    }
    double PEp[7], PNm[7];  // east flux of row p, north flux of row p-1
#pragma unroll
    for (int k = 0; k < 7; ++k) PEp[k] = PNm[k] = 0.0;

    for (int p = r0 - 1; p < r1; ++p) {
        asm volatile("" ::: "memory");
        int e = lane;
        asm volatile("" : "+v"(e));
        RowD R;  // row p+1
        {
            const Q7 q_hi = load(p + 2);
            const Q7 q_lo = get_q(p, e), q_mid = get_q(p + 1, e);
            put_q(p + 2, q_hi);  // takes row p's slot, read above
            const unsigned cl = row_data(q_lo, q_mid, q_hi, p + 1, R);
            count_pred(cl, p + 1);
        }
        const bool lp = live(p), lp1 = live(p + 1);
        // row p: east, south and west faces (the order the reference sums them)
        double acc[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const double west = from_prev_lane(PEp[k]);
            double a = 0.0;
            a = a + PEp[k] * dy;
            a = a + (-PNm[k]) * dx;
            a = a + (-west) * dy;
            acc[k] = a;
        }
        // east face of row p+1 (the west face of the next lane's cell)
        if (p + 1 < r1) {
            double Wn[7];
#pragma unroll
            for (int k = 0; k < 7; ++k) Wn[k] = from_next_lane(R.W[k]);
            const double te = face(R.E, Wn, true, PEp);
            if (lp1 || lane_next(lp1)) tmin = tmin <= te ? tmin : te;
        }
        // north face of row p (the south face of row p+1)
        double PN[7];
        {
            asm volatile("" ::: "memory");
            int e2 = lane;
            asm volatile("" : "+v"(e2));
            double Np[7];
#pragma unroll
            for (int k = 0; k < 7; ++k) Np[k] = rowp[1][k][e2];
            const double tn = face(Np, R.S, false, PN);
            if (lp || lp1) tmin = tmin <= tn ? tmin : tn;
        }
        asm volatile("" ::: "memory");
        int e3 = lane;
        asm volatile("" : "+v"(e3));
        if (p >= r0 && own_col) {
            const long long k = (long long)p * nx + c;
            if (!frozen(p)) {
                double mp[7];
#pragma unroll
                for (int f = 0; f < 7; ++f) mp[f] = rowp[0][f][e3];
                GkrSums S;
                S.PSx = acc[0] + PN[0] * dx;
                S.PSy = acc[1] + PN[1] * dx;
                S.PUx = acc[2] + PN[2] * dx;
                S.PUy = acc[3] + PN[3] * dx;
                S.PBx = acc[4] + PN[4] * dx;
                S.PBy = acc[5] + PN[5] * dx;
                S.PBz = acc[6] + PN[6] * dx;
                unsigned clamps = 0;
                const Q7 Q2 = gkr_update(A, mp, S, h, A.y0 + p * dy, clamps);
                gkr_count_clamps(A, clamps, 1);
                A.s2[k] = make_double2(Q2.v[0], Q2.v[1]);
                A.u2[k] = make_double2(Q2.v[2], Q2.v[3]);
                A.b2[k] = make_double3(Q2.v[4], Q2.v[5], Q2.v[6]);
            } else {
                // kernel.cu:2688-2695: frozen cells keep their state
                A.s2[k] = A.s[k];
                A.u2[k] = A.u[k];
                A.b2[k] = A.b[k];
            }
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            PNm[k] = PN[k];
            rowp[0][k][e3] = R.m[k];
            rowp[1][k][e3] = R.N[k];
        }
    }
    const double w = wave_min(tmin);
    if (lane == 0 && w < 1e300) atomic_min_nonneg(A.tmin_bits, w);
}

}  // namespace gd2