// gd2_gkr.hpp -- the Godunov-Kolgan-Rodionov predictor-corrector MHD step
// (Kernel_GKR, kernel.cu:2663-3284, with predictor kernel.cu:1611-1720 and
// POTOK_Korolkov Gas_Dinamic.cu:70-181) in the reference AoS layout.
//
// One thread per cell, as the reference evaluates it: the cell gathers its
// 13-point stencil (radius 2 along the axes plus the four diagonals) through
// the step's own side rules, forms minmod slopes of itself and its four
// neighbours, a half-step predictor of each of the five (physical fluxes of
// the half-slope face states plus the axisymmetric sources), and updates
// itself from the mean of its state and its predictor with the HLLD (or
// HLLC-MHD) fluxes of the predictor-averaged face states, over dt/2 with the
// axisymmetric sources. Each face is evaluated from the cell's own side with
// its outward normal, as in the reference (the MHD fluxes are not continuous
// at S_M = 0 when the normal field jumps). Stencil points are read from
// memory when needed (served by the caches) rather than held in registers.
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
    double idV;  // 1 / (dx dy)
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

// predictor (kernel.cu:1611-1720): a step of st from the physical fluxes of
// the four half-slope face states (rho, p falling back to the cell value
// when not positive) and the axisymmetric source terms at radius y.
__device__ inline Q7 gkr_predictor(const GkrArgs& A, const Q7& Q, const Q7& DX, const Q7& DY, double st, double y) {
    const double dx = A.dx, dy = A.dy;
    double PSx = 0, PSy = 0, PUx = 0, PUy = 0, PBx = 0, PBy = 0, PBz = 0;
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
        PSx = PSx + P[0] * len;
        PSy = PSy + P[7] * len;
        PUx = PUx + P[1] * len;
        PUy = PUy + P[2] * len;
        PBx = PBx + P[4] * len;
        PBy = PBy + P[5] * len;
        PBz = PBz + P[6] * len;
    }
    // divisions by dV, y and r as products with their reciprocals
    const double sV = st * A.idV, sy = st * fast_rcp(y);
    const double* q = Q.v;
    Q7 QQ;
    double r = q[0] - sV * PSx - sy * q[0] * q[3];
    if (r <= 0) {
        if (A.diag) atomicAdd(&A.diag->clamp_rho, 1ull);
        r = q[0];
    }
    const double ir = fast_rcp(r);
    QQ.v[0] = r;
    QQ.v[2] = (q[0] * q[2] - sV * (PUx) - sy * (q[0] * q[2] * q[3] - q[4] * q[5] * (1.0 / kCpi4))) * ir;
    QQ.v[3] = (q[0] * q[3] - sV * (PUy) - sy * (q[0] * q[3] * q[3] + (q[6] * q[6] - q[5] * q[5]) * (1.0 / kCpi4))) * ir;
    QQ.v[4] = (q[4] - sV * (PBx) - sy * (q[3] * q[4] - q[5] * q[2]));
    QQ.v[5] = (q[5] - sV * (PBy));
    QQ.v[6] = (q[6] - sV * (PBz));
    const double u8 = gkr_u8(A.inv_g1, q[0], q[1], q[2], q[3], q[4], q[5], q[6]);
    double p = (u8 - sV * (PSy) -
                sy * (((u8 + q[1] + (q[4] * q[4] + q[5] * q[5] + q[6] * q[6]) * (1.0 / kCpi8)) * q[3] -
                       q[5] * (q[2] * q[4] + q[3] * q[5]) * (1.0 / kCpi4))) -
                0.5 * r * (QQ.v[2] * QQ.v[2] + QQ.v[3] * QQ.v[3]) -
                (QQ.v[4] * QQ.v[4] + QQ.v[5] * QQ.v[5] + QQ.v[6] * QQ.v[6]) * (1.0 / kCpi8)) * (A.gamma - 1.0);
    if (p <= 0) {
        if (A.diag) atomicAdd(&A.diag->clamp_p, 1ull);
        p = 0.000001;
    }
    QQ.v[1] = p;
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

template <int METOD>
__global__ __launch_bounds__(256) void gkr_kernel(const GkrArgs A) {
    const long long cells = (long long)A.nx * A.ny;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double tmin = 1e300;
    if (idx < cells) {
        const int n = (int)(idx % A.nx), m = (int)(idx / A.nx);
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
            const double hV = h * A.idV, hy = h * fast_rcp(y);
            double r2 = Qb[0] - hV * PSx - hy * Qb[0] * Qb[3];
            if (r2 <= 0) {
                if (A.diag) atomicAdd(&A.diag->clamp_rho, 1ull);
                r2 = Qb[0];
            }
            const double ir2 = fast_rcp(r2);
            const double vx2 =
                (Qb[0] * Qb[2] - hV * (PUx) - hy * (Qb[0] * Qb[2] * Qb[3] - Qb[4] * Qb[5] * (1.0 / kCpi4))) * ir2;
            const double vy2 = (Qb[0] * Qb[3] - hV * (PUy) -
                                hy * (Qb[0] * Qb[3] * Qb[3] + (Qb[6] * Qb[6] - Qb[5] * Qb[5]) * (1.0 / kCpi4))) *
                               ir2;
            const double bx2 = (Qb[4] - hV * (PBx) - hy * (Qb[3] * Qb[4] - Qb[5] * Qb[2]));
            const double by2 = (Qb[5] - hV * (PBy));
            const double bz2 = (Qb[6] - hV * (PBz));
            const double u8 = gkr_u8(A.inv_g1, Qb[0], Qb[1], Qb[2], Qb[3], Qb[4], Qb[5], Qb[6]);
            double p2 = (u8 - hV * (PSy) -
                         hy * (((u8 + Qb[1] + (Qb[4] * Qb[4] + Qb[5] * Qb[5] + Qb[6] * Qb[6]) * (1.0 / kCpi8)) * Qb[3] -
                               Qb[5] * (Qb[2] * Qb[4] + Qb[3] * Qb[5]) * (1.0 / kCpi4))) -
                         0.5 * r2 * (vx2 * vx2 + vy2 * vy2) - (bx2 * bx2 + by2 * by2 + bz2 * bz2) * (1.0 / kCpi8)) *
                        (A.gamma - 1.0);
            if (p2 <= 0) {
                if (A.diag) atomicAdd(&A.diag->clamp_p, 1ull);
                p2 = 0.000001;
            }
            A.s2[idx] = make_double2(r2, p2);
            A.u2[idx] = make_double2(vx2, vy2);
            A.b2[idx] = make_double3(bx2, by2, bz2);
        }
    }
    // exact device-wide minimum of the face time steps (the reference's
    // racy *T fold, kernel.cu:3241-3245)
    const double w = wave_min(tmin);
    if ((threadIdx.x & 63) == 0 && w < 1e300) atomic_min_nonneg(A.tmin_bits, w);
}

}  // namespace gd2
