// gd2_device.hpp -- device-side building blocks of the gfx950 step kernels:
// boundary index mapping, minmod, primitive/conserved conversions, the HLLC
// and exact-Riemann Euler interface fluxes, and wave-level reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gd2.h"

namespace gd2 {

constexpr int kWave = 64;
constexpr double kCpi4 = 12.56637061435917295384;  // 4 pi    (Header.h:20)
constexpr double kCpi8 = 25.13274122871834590768;  // 8 pi    (Header.h:21)
constexpr double kSpi4 = 3.544907701811032;        // sqrt(4 pi) literal (Header.h:22)

// Device counters (clamps, iteration caps) read by gd2_get_diagnostics.
struct DeviceDiag {
    unsigned long long clamp_rho, clamp_p, newton_cap, bisect_cap;
};

// Device clock of a solver context: the next step's dt is cfl / w[cur]
// (w = wave-speed maximum of the current state, accumulated as double bits).
// Device clock of a native solver. Step n takes dt = cfl / w[n % 3] (the
// wave-speed bound of U^n) and folds the bound of U^{n+1} into w[(n+1) % 3];
// w[(n+1) % 3] is zero when step n starts (three slots: a step clears the
// slot of the step after it, so no clearing races with a fold).
struct Clock {
    double t, dt, cfl, pad;
    unsigned long long w[3];
    long long step;
};

__device__ __forceinline__ int clock_next(int cur) { return cur == 2 ? 0 : cur + 1; }

__device__ __forceinline__ double clock_dt(const Clock* c, int cur) {
    return c->cfl / __longlong_as_double((long long)c->w[cur]);
}

// The step's clock update, done by one thread before any wave of the step
// folds into w[(cur+1) % 3]: dt (clipped to t_end when t_end > 0; 0 and no
// step once t reached it), t, the step count, and the slots of the next
// steps cleared.
__device__ __forceinline__ void clock_begin(Clock* c, int cur, double t_end) {
    double dt = clock_dt(c, cur);
    bool live = true;
    if (t_end > 0.0) {
        if (!(c->t < t_end)) {
            dt = 0.0;
            live = false;
        } else if (dt > t_end - c->t) {
            dt = t_end - c->t;
        }
    }
    c->dt = dt;
    c->w[clock_next(cur)] = 0ull;
    c->w[clock_next(clock_next(cur))] = 0ull;
    if (live) {
        c->t = c->t + dt;
        c->step += 1;
    }
}

// The clock update of a step whose kernel computes dt itself (no end time):
// every wave reads only cfl and w[cur]; one thread of the launch writes dt,
// t, the step count and clears w[(cur+2) % 3] (w[(cur+1) % 3], which the
// waves fold into, was cleared by the step before).
__device__ __forceinline__ void clock_commit(Clock* c, int cur, double* dt_log, long long cap) {
    const double dt = clock_dt(c, cur);
    const long long n = c->step;
    c->dt = dt;
    c->t = c->t + dt;
    c->step = n + 1;
    c->w[clock_next(clock_next(cur))] = 0ull;
    if (dt_log && n < cap) dt_log[n] = dt;
}

// Logical index k (may be -2..n+1, or further on padding lanes) -> stored
// index; flip is set where the reflective rule negates the normal vector
// components. Transmissive and frozen sides clamp (kernel.cu:1762-1793,
// 3374-3417), reflective sides mirror (kernel.cu:2263-2277, 2356-2405),
// periodic sides wrap.
__device__ __forceinline__ int map_index(int k, int n, int lo, int hi, bool& flip) {
    flip = false;
    if (k < 0) {
        if (lo == GD2_BC_PERIODIC) return ((k % n) + n) % n;
        if (lo == GD2_BC_REFLECTIVE) {
            flip = true;
            int m = -1 - k;
            return m < n ? m : n - 1;
        }
        return 0;
    }
    if (k >= n) {
        if (hi == GD2_BC_PERIODIC) return k % n;
        if (hi == GD2_BC_REFLECTIVE) {
            flip = true;
            int m = 2 * n - 1 - k;
            return m >= 0 ? m : 0;
        }
        return n - 1;
    }
    return k;
}

// minmod(a, b) of kernel.cu:51-62 written as the median of (0, a, b): equal
// for every input including signed zeros' magnitudes (sign-sum zero -> 0,
// same signs -> the argument of smaller magnitude).
__device__ __forceinline__ double minmod(double a, double b) {
    return fmax(fmin(a, b), fmin(fmax(a, b), 0.0));
}

// Reciprocal and square root from the hardware approximations (v_rcp_f64,
// v_rsq_f64, relative error about 2^-24.6) refined by Newton / Goldschmidt
// steps, at a fraction of the instruction count of the IEEE division and
// sqrt expansions. For finite positive normal arguments. fast_rcp and the
// sound speeds take one step (measured over 4e6 arguments: at most 10 and 19
// units of 2^-52 from the IEEE results, i.e. relative 2e-15 and 4e-15, three
// orders below the 1e-12 per-step parity bound); fast_sqrt adds a final
// correction of s.
__device__ __forceinline__ double fast_rcp(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double fast_sqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, h = 0.5 * y;
    const double r = fma(-s, h, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    const double d = fma(-s, s, x);
    s = fma(d, h, s);
    return x == 0.0 ? 0.0 : s;
}

// sqrt(a / b) for a, b > 0 as 2a * (1 / (2 sqrt(a b))): one reciprocal square
// root refined by one Goldschmidt step (h -> 1/(2 sqrt(x))), instead of a
// reciprocal and a square root. Returns 0 when a b == 0.
// GUARD = false: the caller guarantees a b > 0 (normal).
template <bool GUARD = true>
__device__ __forceinline__ double sqrt_ratio(double a, double b) {
    const double x = a * b;
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y;
    double h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    h = fma(h, r, h);
    if constexpr (GUARD)
        return x == 0.0 ? 0.0 : (a + a) * h;
    else
        return (a + a) * h;
}

// A wave-uniform pointer pinned to scalar registers. Row base pointers pass
// through this so that loop strength reduction cannot turn them into per-lane
// 64-bit induction variables: loads then take the scalar-base + 32-bit
// vector-offset form.
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<T*>(((unsigned long long)hi << 32) | lo);
}

// Raw buffer resource over a wave-uniform base pointer (scalar registers):
// accesses then take a 32-bit per-lane byte offset, with no 64-bit address
// arithmetic per lane; offsets at or beyond 2^31 - 1 read 0 / drop the store.
constexpr int kBufFlags = 0x00020000;  // raw dword access on CDNA3/CDNA4

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7fffffff, kBufFlags);
}

__device__ __forceinline__ double ld_buf(__amdgpu_buffer_rsrc_t r, unsigned byte_off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0));
}

__device__ __forceinline__ void st_buf(__amdgpu_buffer_rsrc_t r, unsigned byte_off, double v) {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, byte_off, 0, 0);
}

// Value held by lane + 1 (lane 63 receives 0): DPP wave_shl:1 on both halves.
__device__ __forceinline__ double from_next_lane(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x130, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Value held by lane - 1 (lane 0 receives 0): DPP wave_shr:1 on both halves.
__device__ __forceinline__ double from_prev_lane(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sound speed c = sqrt(a / b) (a = gamma p, b = rho) with h = 1 / (2 sqrt(a b)):
// dc/dp = gamma h and dc/drho = -c^2 h, so the derivatives need no division.
struct SoundSpeed {
    double c, h;
};

// The same steps as sqrt_ratio (the adjoint recomputes the forward's c
// bit for bit).
template <bool GUARD = true>
__device__ __forceinline__ SoundSpeed sound_speed(double a, double b) {
    const double x = a * b;
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y;
    double h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    h = fma(h, r, h);
    SoundSpeed s;
    if constexpr (GUARD) {
        s.c = x == 0.0 ? 0.0 : (a + a) * h;
        s.h = x == 0.0 ? 0.0 : h;
    } else {
        s.c = (a + a) * h;
        s.h = h;
    }
    return s;
}

// Primitive Euler state.
struct PrimE {
    double r, u, v, p;
};

__device__ __forceinline__ PrimE cons_to_prim_e(double r, double mx, double my, double E, double g1) {
    PrimE q;
    const double inv = fast_rcp(r);
    q.r = r;
    q.u = mx * inv;
    q.v = my * inv;
    q.p = g1 * (E - 0.5 * (mx * q.u + my * q.v));
    return q;
}

// Face flux in the face frame: (mass, normal momentum, tangential momentum,
// energy) along the face normal.
struct FluxE {
    double fr, fn, ft, fe;
};

// HLLC flux of HLLC_Aleksashov (kernel.cu:853-1165) with B = 0, in its
// face frame: normal velocity un, tangential ut. Only the star state of the
// side chosen by sign(SM) is formed (the reference forms both); the branch
// choice SL > 0 / SL <= 0 <= SM / SM <= 0 <= SR / SR < 0 (kernel.cu:1119-1158)
// becomes a select. smax returns max(|SL|, |SR|). With Q, *pq receives the
// flux of a passive scalar Q carried with the density, as HLLCQ_Aleksashov
// forms it with B = 0 (Gas_Dinamic.cu:740-772).
template <bool GUARD = true, bool Q = false>
__device__ __forceinline__ FluxE hllc_euler(double rL, double unL, double utL, double pL,
                                            double rR, double unR, double utR, double pR,
                                            double gamma, double inv_g1, double& smax,
                                            double QL = 0.0, double QR = 0.0, double* pq = nullptr) {
    const double cL = sqrt_ratio<GUARD>(gamma * pL, rL);
    const double cR = sqrt_ratio<GUARD>(gamma * pR, rR);
    const double SL = fmin(unL - cL, unR - cR);
    const double SR = fmax(unL + cL, unR + cR);
    smax = fmax(fabs(SL), fabs(SR));
    const double suL = SL - unL, suR = SR - unR;
    const double mL = suL * rL, mR = suR * rR;
    const double smR = mR * unR;  // separately rounded products, as the reference
    const double smL = mL * unL;
    const double SM = (smR - pR + pL - smL) * fast_rcp(mR - mL);
    // averaged star total pressure (kernel.cu:1069-1071)
    const double ptz = 0.5 * ((pR + mR * (SM - unR)) + (pL + mL * (SM - unL)));
    const bool left = SM >= 0.0;
    const double S = left ? SL : SR;
    const double r = left ? rL : rR;
    const double un = left ? unL : unR;
    const double ut = left ? utL : utR;
    const double p = left ? pL : pR;
    const double su = S - un;  // = suL or suR, the same operation on the same operands
    const bool super = (SL > 0.0) || (SR < 0.0);
    const double e = p * inv_g1 + r * (0.5 * (un * un + ut * ut));
    const double den = super ? 1.0 : (S - SM);
    const double inv = fast_rcp(den);
    const double sum = su * inv;
    const double rz = r * sum;
    const double ez = e * sum + (ptz * SM - p * un) * inv;
    const double Se = super ? 0.0 : S;
    const double run = r * un;
    FluxE F;
    F.fr = run + Se * (rz - r);
    F.fn = run * un + p + Se * (SM * rz - run);
    // the reference's run ut + S (ut rz - r ut) is ut times the mass flux
    // (the tangential velocity is carried across the contact unchanged)
    F.ft = ut * F.fr;
    F.fe = (e + p) * un + Se * (ez - e);
    if constexpr (Q) {
        const double q = left ? QL : QR;
        *pq = q * un + Se * (rz * q / r - q);
    }
    return F;
}

// HLL flux of kernel.cu:1485-1609 in the face frame: wave-speed bounds
// DL = min(unL, unR) - max(cL, cR), DR = max(unL, unR) + max(cL, cR) (sound
// speed 0 where rho <= 0); the left flux when DL > 0, the right one when
// DR < 0, else the HLL average (DR FL - DL FR + DL DR (UR - UL)) / (DR - DL).
// smax returns max(|DL|, |DR|). GUARD = false: positive densities.
template <bool GUARD = true>
__device__ __forceinline__ FluxE hll_euler(double rL, double unL, double utL, double pL,
                                           double rR, double unR, double utR, double pR,
                                           double gamma, double inv_g1, double& smax) {
    double cL, cR;
    if constexpr (GUARD) {
        cL = rL <= 0.0 ? 0.0 : sqrt_ratio(gamma * pL, rL);
        cR = rR <= 0.0 ? 0.0 : sqrt_ratio(gamma * pR, rR);
    } else {
        cL = sqrt_ratio<false>(gamma * pL, rL);
        cR = sqrt_ratio<false>(gamma * pR, rR);
    }
    const double cm = fmax(cL, cR);
    const double DL = fmin(unL, unR) - cm;
    const double DR = fmax(unL, unR) + cm;
    smax = fmax(fabs(DL), fabs(DR));
    const double eL = pL * inv_g1 + rL * (0.5 * (unL * unL + utL * utL));
    const double eR = pR * inv_g1 + rR * (0.5 * (unR * unR + utR * utR));
    const double mL = rL * unL, mR = rR * unR;
    FluxE FL, FR;
    FL.fr = mL;
    FL.fn = mL * unL + pL;
    FL.ft = mL * utL;
    FL.fe = (eL + pL) * unL;
    FR.fr = mR;
    FR.fn = mR * unR + pR;
    FR.ft = mR * utR;
    FR.fe = (eR + pR) * unR;
    const double den = DR - DL;
    const double inv = den > 0.0 ? fast_rcp(den) : 0.0;  // den = 0 only for two states at rest with c = 0
    const double dd = DL * DR;
    FluxE F;
    F.fr = (DR * FL.fr - DL * FR.fr + dd * (rR - rL)) * inv;
    F.fn = (DR * FL.fn - DL * FR.fn + dd * (mR - mL)) * inv;
    F.ft = (DR * FL.ft - DL * FR.ft + dd * (rR * utR - rL * utL)) * inv;
    F.fe = (DR * FL.fe - DL * FR.fe + dd * (eR - eL)) * inv;
    const bool left = DL > 0.0, right = DR < 0.0;
    F.fr = left ? FL.fr : (right ? FR.fr : F.fr);
    F.fn = left ? FL.fn : (right ? FR.fn : F.fn);
    F.ft = left ? FL.ft : (right ? FR.ft : F.ft);
    F.fe = left ? FL.fe : (right ? FR.fe : F.fe);
    return F;
}

// HLLC_Korolkov (kernel.cu:1349-1483) in the face frame: the HLL solver's
// outer speeds D_L = min(unL, unR) - max(cL, cR), D_R = max(unL, unR) +
// max(cL, cR); contact speed D_C and star pressure P_T of the two-state
// HLLC; the star state takes normal velocity D_C, keeps the tangential one
// and takes its energy from the equation of state, P_T/(gamma-1) +
// rho* (D_C^2 + ut^2)/2 (kernel.cu:1448, 1468). The reference's lab-frame
// projections become face-frame components (a rotation: the same flux).
// Branches D_L > 0 / D_R < 0 / D_C >= 0 (kernel.cu:1399-1480) are selects.
template <bool GUARD = true>
__device__ __forceinline__ FluxE hllc_korolkov_euler(double rL, double unL, double utL, double pL,
                                                     double rR, double unR, double utR, double pR,
                                                     double gamma, double inv_g1, double& smax) {
    const double cL = sqrt_ratio<GUARD>(gamma * pL, rL);
    const double cR = sqrt_ratio<GUARD>(gamma * pR, rR);
    const double cm = fmax(cL, cR);
    const double DL = fmin(unL, unR) - cm;
    const double DR = fmax(unL, unR) + cm;
    smax = fmax(fabs(DL), fabs(DR));
    const double mL = (DL - unL) * rL, mR = (DR - unR) * rR;
    const double smR = mR * unR, smL = mL * unL;
    const double idm = fast_rcp(mR - mL);
    const double DC = (smR - smL - pR + pL) * idm;
    // P_T = (pL mR - pR mL + mL mR (unR - unL)) / (mR - mL)   (kernel.cu:1441)
    const double PT = (pL * mR - pR * mL + (mL * mR) * (unR - unL)) * idm;
    const bool left = DC >= 0.0;
    const double S = left ? DL : DR;
    const double r = left ? rL : rR;
    const double un = left ? unL : unR;
    const double ut = left ? utL : utR;
    const double p = left ? pL : pR;
    const bool super = (DL > 0.0) || (DR < 0.0);
    const bool sup_left = DL > 0.0;
    // a supersonic face takes the upwind state's physical flux
    const double r2 = super ? (sup_left ? rL : rR) : r;
    const double un2 = super ? (sup_left ? unL : unR) : un;
    const double ut2 = super ? (sup_left ? utL : utR) : ut;
    const double p2 = super ? (sup_left ? pL : pR) : p;
    const double e = p2 * inv_g1 + r2 * (0.5 * (un2 * un2 + ut2 * ut2));
    const double den = super ? 1.0 : (S - DC);
    const double rz = r * ((S - un) * fast_rcp(den));
    const double ez = PT * inv_g1 + rz * (0.5 * (DC * DC + ut * ut));
    const double Se = super ? 0.0 : S;
    const double run = r2 * un2;
    FluxE F;
    F.fr = run + Se * (rz - r2);
    F.fn = run * un2 + p2 + Se * (rz * DC - run);
    F.ft = run * ut2 + Se * (rz * ut2 - r2 * ut2);
    F.fe = (e + p2) * un2 + Se * (ez - e);
    return F;
}

// A reflective west or south wall face evaluated as F(ghost | interior) with
// the face normal along +x / +y. HLLC_Korolkov's star energy (equation of
// state, not the Rankine-Hugoniot condition) makes its flux jump at D_C = 0,
// where a mirrored pair puts the wall face, and the reference evaluates that
// face from the interior cell with the outward normal (kernel.cu:1803-1834,
// 3554-3587): by the mirror symmetry of the pair, that is the flux above with
// the mass, tangential momentum and energy negated. The other fluxes are
// continuous at their contact, where both evaluations agree.
template <int FLUX>
__device__ __forceinline__ void wall_face_fix(FluxE& F, bool wall) {
    if constexpr (FLUX == GD2_FLUX_HLLC_KOROLKOV) {
        if (wall) {
            F.fr = -F.fr;
            F.ft = -F.ft;
            F.fe = -F.fe;
        }
    }
}

// HLLC_Korolkov_2D / HLLCQ_Korolkov_2D (Gas_Dinamic.cu:265-464) in the face
// frame: Toro's HLLC with S_L = min(uL - cL, uR - cR), S_R = max(uL + cL,
// uR + cR), the one-sided star pressure pp = pL + rhoL (S_L - uL)(S_M - uL),
// the star energy ((S - u) e - p u + pp S_M) / (S - S_M), and S_M <= 0 on the
// right star state (Gas_Dinamic.cu:433). With Q, *pq receives the flux of the
// passive scalar Q (carried with the density; Gas_Dinamic.cu:413-455).
template <bool GUARD = true, bool Q = false>
__device__ __forceinline__ FluxE hllc_toro_euler(double rL, double unL, double utL, double pL,
                                                 double rR, double unR, double utR, double pR,
                                                 double gamma, double inv_g1, double& smax,
                                                 double QL = 0.0, double QR = 0.0, double* pq = nullptr) {
    const double cL = sqrt_ratio<GUARD>(gamma * pL, rL);
    const double cR = sqrt_ratio<GUARD>(gamma * pR, rR);
    const double SL = fmin(unL - cL, unR - cR);
    const double SR = fmax(unL + cL, unR + cR);
    smax = fmax(fabs(SL), fabs(SR));
    const double mL = (SL - unL) * rL, mR = (SR - unR) * rR;
    const double smR = mR * unR, smL = mL * unL;
    const double SM = (smR - smL - pR + pL) * fast_rcp(mR - mL);
    const double pp = pL + mL * (SM - unL);
    const bool sup_left = SL >= 0.0, sup_right = SR <= 0.0;
    const bool super = sup_left || sup_right;
    const bool left = super ? sup_left : (SM > 0.0);
    const double S = left ? SL : SR;
    const double r = left ? rL : rR;
    const double un = left ? unL : unR;
    const double ut = left ? utL : utR;
    const double p = left ? pL : pR;
    const double e = p * inv_g1 + r * (0.5 * (un * un + ut * ut));
    const double inv = fast_rcp(super ? 1.0 : (S - SM));
    const double su = S - un;
    const double rr = r * su * inv;
    const double ee = (su * e - p * un + pp * SM) * inv;
    const double Se = super ? 0.0 : S;
    const double run = r * un;
    FluxE F;
    F.fr = Se * (rr - r) + run;
    F.fe = Se * (ee - e) + (e + p) * un;
    F.fn = (run * un + p) + Se * (rr * SM - run);
    F.ft = run * ut + Se * (rr * ut - r * ut);
    if constexpr (Q) {
        const double q = left ? QL : QR;
        *pq = Se * (rr * q / r - q) + q * un;
    }
    return F;
}

// HLLC_Aleksashov2 (kernel.cu:1167-1348) in the frame (n, t = (-n2, n1)) of
// its caller's normal: HLLC_Aleksashov's HLLC with B = 0, except that the
// contact speed's numerator takes the left state's tangential velocity in
// the term suL r1 vL[1] (kernel.cu:1222). That breaks the mirror symmetry
// F(L, R, n) = -F(R, L, -n) and the consistency F(U, U) = F(U), so it is
// offered pointwise (gd2_euler_face_fluxes), not as a stage flux.
__device__ __forceinline__ FluxE hllc_aleksashov2_euler(double rL, double unL, double utL, double pL,
                                                        double rR, double unR, double utR, double pR,
                                                        double gamma, double inv_g1, double& smax) {
    const double cL = sqrt(gamma * pL / rL);
    const double cR = sqrt(gamma * pR / rR);
    const double SL = fmin(unL - cL, unR - cR);
    const double SR = fmax(unL + cL, unR + cR);
    smax = fmax(fabs(SL), fabs(SR));
    const double suR = SR - unR, suL = SL - unL;
    const double dd = suR * rR - suL * rL;
    const double SM = fabs(dd) > 0.0 ? (suR * rR * unR - pR + pL - suL * rL * utL) / dd : 0.0;
    const double ptz = ((pR + rR * suR * (SM - unR)) + (pL + rL * suL * (SM - unL))) / 2.0;
    const bool left = SM >= 0.0;
    const double S = left ? SL : SR;
    const double r = left ? rL : rR;
    const double un = left ? unL : unR;
    const double ut = left ? utL : utR;
    const double p = left ? pL : pR;
    const double su = left ? suL : suR;
    // branch order of kernel.cu:1301-1341: SL > 0; SL <= 0 <= SM (left star,
    // whatever SR is: this SM can lie outside [SL, SR]); SM <= 0 <= SR; SR < 0
    const bool sup_left = SL > 0.0, sup_right = !sup_left && !left && SR < 0.0;
    const bool super = sup_left || sup_right;
    const double r2 = super ? (sup_left ? rL : rR) : r;
    const double un2 = super ? (sup_left ? unL : unR) : un;
    const double ut2 = super ? (sup_left ? utL : utR) : ut;
    const double p2 = super ? (sup_left ? pL : pR) : p;
    const double e = p2 * inv_g1 + r2 * (0.5 * (un2 * un2 + ut2 * ut2));
    const double den = super ? 1.0 : (S - SM);
    const double sm = su / den;
    const double rz = r * sm;
    const double ez = e * sm + (ptz * SM - p * un) / den;
    const double Se = super ? 0.0 : S;
    FluxE F;
    F.fr = r2 * un2 + Se * (rz - r2);
    F.fe = (e + p2) * un2 + Se * (ez - e);
    F.fn = r2 * un2 * un2 + p2 + Se * (SM * rz - r2 * un2);
    F.ft = r2 * un2 * ut2 + Se * (ut2 * rz - r2 * ut2);
    return F;
}

// ---------------------------------------------------------------------------
// Exact Riemann solver (Godunov_Solver_Alexashov, kernel.cu:214-656; newton
// 721-766; devtwo 768-835; lev 837-851) with gamma at run time. The
// reference loops without a bound; the caps below are never reached on
// finite inputs and each hit is counted.
// ---------------------------------------------------------------------------
constexpr int kNewtonCap = 200;
constexpr int kBisectCap = 400;

struct GConst {
    double ga, g1, g2, gp, gm;
};

__device__ __forceinline__ GConst make_gconst(double gamma) {
    GConst k;
    k.ga = gamma;
    k.g1 = gamma - 1.0;
    k.g2 = gamma + 1.0;
    k.gp = (k.g2 / k.ga) / 2.0;
    k.gm = (k.g1 / k.ga) / 2.0;
    return k;
}

__device__ inline double lev_f(const GConst& k, double enI, double pI, double rI, double enII, double pII,
                               double rII, double uuu) {
    double cI = sqrt(k.ga * pI / rI);
    double cII = sqrt(k.ga * pII / rII);
    double fI = (uuu - pI) / (rI * cI * sqrt(k.gp * uuu / pI + k.gm));
    double fII = 2.0 / k.g1 * cII * (pow((uuu / pII), k.gm) - 1.0);
    return (fI + fII) - (enI - enII);
}

__device__ inline double devtwo(const GConst& k, double enI, double pI, double rI, double enII, double pII,
                                double rII, DeviceDiag* diag) {
    const double epsil = 10e-10;
    double kl = pI, kp = pII;
    double ksi = lev_f(k, enI, pI, rI, enII, pII, rII, kl);
    double ksir = lev_f(k, enI, pI, rI, enII, pII, rII, kp);
    if (fabs(ksi) <= epsil) return kl;
    if (fabs(ksir) <= epsil) return kp;
    for (int it = 1;; ++it) {
        double kc = (kl + kp) / 2.0;
        double ksit = lev_f(k, enI, pI, rI, enII, pII, rII, kc);
        if (fabs(ksit) <= epsil) return kc;
        if (it >= kBisectCap) {
            if (diag) atomicAdd(&diag->bisect_cap, 1ull);
            return kc;
        }
        if ((ksi * ksit) <= 0.0) {
            kp = kc;
        } else {
            kl = kc;
            ksi = ksit;
        }
    }
}

__device__ inline double newton(const GConst& k, double enI, double pI, double rI, double enII, double pII,
                                double rII, DeviceDiag* diag) {
    double cI = sqrt(k.ga * pI / rI);
    double cII = sqrt(k.ga * pII / rII);
    double pn = pI * rII * cII + pII * rI * cI + (enI - enII) * rI * cI * rII * cII;
    pn = pn / (rI * cI + rII * cII);
    const double pee = pn;
    for (int kiter = 1;; ++kiter) {
        double p = pn;
        double fI = (p - pI) / (rI * cI * sqrt(k.gp * p / pI + k.gm));
        double fIs = (k.ga + 1.0) * p / pI + (3.0 * k.ga - 1.0);
        fIs = fIs / (4.0 * k.ga * rI * cI * pow((k.gp * p / pI + k.gm), (3.0 / 2.0)));
        double fII = (p - pII) / (rII * cII * sqrt(k.gp * p / pII + k.gm));
        double fIIs = (k.ga + 1.0) * p / pII + (3.0 * k.ga - 1.0);
        fIIs = fIIs / (4.0 * k.ga * rII * cII * pow((k.gp * p / pII + k.gm), (3.0 / 2.0)));
        pn = p - (fI + fII - (enI - enII)) / (fIs + fIIs);
        if (!(fabs(pn / pee - p / pee) >= 1e-10)) break;
        if (kiter >= kNewtonCap) {
            if (diag) atomicAdd(&diag->newton_cap, 1ull);
            break;
        }
    }
    return pn;
}

__device__ __forceinline__ double fan_a(const GConst& k, double p, double pK, double rK, double cK) {
    if (fabs(p - pK) < 1e-10) return rK * cK;
    return k.gm * rK * cK * (1.0 - p / pK) / (1.0 - pow((p / pK), k.gm));
}

// Exact-Riemann flux in the face frame (normal along +n, tangential ut).
// Returns the face time step numerator factor: UU (the largest wave-speed
// magnitude, clamped below by 1e-8 as in kernel.cu:331-339) through smax.
__device__ inline FluxE exact_euler(double rL, double unL, double utL, double pL,
                                    double rR, double unR, double utR, double pR,
                                    double gamma, double& smax, DeviceDiag* diag) {
    const GConst k = make_gconst(gamma);
    double enI = unL, teI = utL, pI = pL, rI = rL;
    double enII = unR, teII = utR, pII = pR, rII = rR;
    bool ipiz = false;
    if (pI > pII) {  // kernel.cu:243-270
        enI = -unR; teI = utR; pI = pR; rI = rR;
        enII = -unL; teII = utL; pII = pL; rII = rL;
        ipiz = true;
    }
    double cI = 0.0, cII = 0.0;
    if (rI != 0.0) cI = sqrt(k.ga * pI / rI);
    if (rII != 0.0) cII = sqrt(k.ga * pII / rII);
    const double a = sqrt(rI * (k.g2 * pII + k.g1 * pI) / 2.0);
    const double Uud = (pII - pI) / a;
    const double Urz = -2.0 * cII / k.g1 * (1.0 - pow((pI / pII), k.gm));
    const double Uvk = -2.0 * (cII + cI) / k.g1;
    const double Udf = enI - enII;
    int il = -2, ip = -2;
    double p = 0.0;
    if (Udf < Uvk) {
        il = -1; ip = -1;
    } else if ((Udf >= Uvk) && (Udf <= Urz)) {
        p = pI * pow(((Udf - Uvk) / (Urz - Uvk)), (1.0 / k.gm));
        il = 0; ip = 0;
    } else if ((Udf > Urz) && (Udf <= Uud)) {
        p = devtwo(k, enI, pI, rI, enII, pII, rII, diag);
        il = 1; ip = 0;
    } else if (Udf > Uud) {
        p = newton(k, enI, pI, rI, enII, pII, rII, diag);
        il = 1; ip = 1;
    }
    const double w = 0.0;
    double r = 0.0, te = 0.0, en = 0.0, UU = 0.0;
    if (il == 1 && ip == 1) {  // two shocks, kernel.cu:320-374
        double aI = sqrt(rI * (k.g2 / 2.0 * p + k.g1 / 2.0 * pI));
        double aII = sqrt(rII * (k.g2 / 2.0 * p + k.g1 / 2.0 * pII));
        double u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        double dI = enI - aI / rI;
        double dII = enII + aII / rII;
        UU = fmax(fabs(dI), fabs(dII));
        if (w <= dI) { en = enI; p = pI; r = rI; te = teI; }
        else if ((w > dI) && (w <= u)) { en = u; r = rI * aI / (aI - rI * (enI - u)); te = teI; }
        else if ((w > u) && (w < dII)) { en = u; r = rII * aII / (aII + rII * (enII - u)); te = teII; }
        else if (w >= dII) { en = enII; p = pII; r = rII; te = teII; }
    } else if (il == 1 && ip == 0) {  // shock + fan, kernel.cu:378-450
        double aI = sqrt(rI * (k.g2 / 2.0 * p + k.g1 / 2.0 * pI));
        double aII = fan_a(k, p, pII, rII, cII);
        double u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        double dI = enI - aI / rI;
        double dII = enII + cII;
        double ddII = u + cII - k.g1 * (enII - u) / 2.0;
        UU = fmax(fmax(fabs(dI), fabs(dII)), fabs(ddII));
        if (w <= dI) { en = enI; p = pI; r = rI; te = teI; }
        if ((w > dI) && (w <= u)) { en = u; r = rI * aI / (aI - rI * (enI - u)); te = teI; }
        if ((w > u) && (w <= ddII)) {
            double ce = cII - k.g1 / 2.0 * (enII - u);
            en = u; r = k.ga * p / ce / ce; te = teII;
        }
        if ((w > ddII) && (w < dII)) {
            double ce = -k.g1 / k.g2 * (enII - w) + 2.0 / k.g2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / k.gm)); r = k.ga * p / ce / ce; te = teII;
        }
        if (w >= dII) { en = enII; p = pII; r = rII; te = teII; }
    } else if (il == 0 && ip == 0) {  // two fans, kernel.cu:452-563
        double aI = fan_a(k, p, pI, rI, cI);
        double aII = fan_a(k, p, pII, rII, cII);
        double u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        double dI = enI - cI;
        double ddI = u - cI - k.g1 * (enI - u) / 2.0;
        double dII = enII + cII;
        double ddII = u + cII - k.g1 * (enII - u) / 2.0;
        UU = fmax(fmax(fmax(fabs(dI), fabs(dII)), fabs(ddII)), fabs(ddI));
        if (w <= dI) { en = enI; p = pI; r = rI; te = teI; }
        if ((w > dI) && (w < ddI)) {
            double ce = k.g1 / k.g2 * (enI - w) + 2.0 / k.g2 * cI;
            en = w + ce; p = pI * pow((ce / cI), (1.0 / k.gm)); r = k.ga * p / ce / ce; te = teI;
        }
        if ((w >= ddI) && (w <= u)) {
            double ce = cI + k.g1 / 2.0 * (enI - u);
            en = u; r = k.ga * p / ce / ce; te = teI;
        }
        if ((w > u) && (w <= ddII)) {
            double ce = cII - k.g1 / 2.0 * (enII - u);
            en = u; r = k.ga * p / ce / ce; te = teII;
        }
        if ((w > ddII) && (w < dII)) {
            double ce = -k.g1 / k.g2 * (enII - w) + 2.0 / k.g2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / k.gm)); r = k.ga * p / ce / ce; te = teII;
        }
        if (w >= dII) { en = enII; p = pII; r = rII; te = teII; }
    } else if (il == -1 && ip == -1) {  // vacuum, kernel.cu:566-630
        double dI = enI - cI;
        double ddI = enI + 2.0 / k.g1 * cI;
        double dII = enII + cII;
        double ddII = enII - 2.0 / k.g1 * cII;
        UU = fmax(fmax(fmax(fabs(dI), fabs(dII)), fabs(ddII)), fabs(ddI));
        if (w <= dI) { en = enI; p = pI; r = rI; te = teI; }
        if ((w > dI) && (w < ddI)) {
            double ce = k.g1 / k.g2 * (enI - w) + 2.0 / k.g2 * cI;
            en = w + ce; p = pI * pow((ce / cI), (1.0 / k.gm)); r = k.ga * p / ce / ce; te = teI;
        }
        if ((w >= ddI) && (w <= ddII)) { en = w; p = 0.0; r = 0.0; te = 0.0; }
        if ((w > ddII) && (w < dII)) {
            double ce = -k.g1 / k.g2 * (enII - w) + 2.0 / k.g2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / k.gm)); r = k.ga * p / ce / ce; te = teII;
        }
        if (w >= dII) { en = enII; p = pII; r = rII; te = teII; }
    } else {
        smax = __builtin_nan("");
        FluxE bad = {smax, smax, smax, smax};
        return bad;
    }
    smax = UU > 1e-8 ? UU : 1e-8;
    if (ipiz) en = -en;
    // kernel.cu:639-651 in the face frame (the third frame vector is zero)
    const double eo = p / k.g1 + 0.5 * r * (en * en + te * te);
    FluxE F;
    F.fr = r * en;
    F.fn = r * en * en + p;
    F.ft = r * en * te;
    F.fe = en * eo + en * p;
    return F;
}

// Wave-wide max of a non-negative double (DPP/permute butterfly).
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmin(v, __shfl_xor(v, off, kWave));
    return v;
}

// Bit patterns of non-negative doubles order like the doubles, so a device
// max / min over them is an exact, order-independent reduction.
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* dst, double v) {
    atomicMax(dst, (unsigned long long)__double_as_longlong(v));
}

__device__ __forceinline__ void atomic_min_nonneg(unsigned long long* dst, double v) {
    atomicMin(dst, (unsigned long long)__double_as_longlong(v));
}

// XCD-aware tile order: the hardware hands consecutive workgroups to the 8
// XCDs in turn; remap so that each XCD walks a contiguous range of tiles (its
// neighbours' halo rows and columns then sit in the same L2).
// Issue priority between the waves sharing a SIMD. The older of two waves
// wins issue ties, so one would finish far ahead and leave the other to run
// alone. Every kTurnPeriod passes a wave publishes its count of such turns in
// its SIMD's row of g_simd_passes (slot = hardware wave slot; turns + 1 while
// it runs, 0 when it has none; 16 slots for each of the 8 x 256 x 4 SIMDs)
// and reads the row one turn earlier: the wave that is behind takes the
// higher priority, level waves alternate by turn and slot parity. Measured on
// the one-pass step at 4096^2 (ms per step): a turn every pass 0.554, none
// 0.500, every 2 passes 0.479, every 4 0.463-0.469, every 8 0.467-0.471,
// every 16 0.471, every 32 0.500. Re-deciding priority every pass keeps the
// two waves in step, so they wait for memory at the same time; every few
// passes they interleave and still finish together.
constexpr int kTurnPeriod = 4;
constexpr int kSimdSlots = 16;
// one table per translation unit (waves of one kernel share its table)
static __device__ unsigned g_simd_passes[8 * 256 * 4 * kSimdSlots];

struct SimdTurn {
    unsigned* slots;
    unsigned wid;
    unsigned seen;  // lane l < 16: slot l as read one pass earlier
    unsigned done;  // passes done by this wave

    __device__ __forceinline__ void init() {
        const unsigned hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);        // HW_ID
        const unsigned xcc = __builtin_amdgcn_s_getreg((15 << 11) | 20) & 7u; // XCC_ID
        wid = hw & 15u;
        const unsigned key = (((xcc * 8u + ((hw >> 13) & 7u)) * 2u + ((hw >> 12) & 1u)) * 16u + ((hw >> 8) & 15u)) * 4u +
                             ((hw >> 4) & 3u);
        slots = g_simd_passes + key * kSimdSlots;
        seen = 0u;
        done = 0u;
    }

    __device__ __forceinline__ void step(int pass) {
        const unsigned lane = threadIdx.x & 63u;
        unsigned v = (lane < (unsigned)kSimdSlots && lane != wid) ? seen : 0u;
        // maximum over the 16 lanes of DPP row 0 (row_ror 8, 4, 2, 1)
        v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false));
        v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false));
        v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x122, 0xF, 0xF, false));
        v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x121, 0xF, 0xF, false));
        const unsigned other = __builtin_amdgcn_readfirstlane(v);
        const unsigned mine = done + 1u;
        bool high;
        if (other == 0u)
            high = false;  // alone on the SIMD
        else if (other != mine)
            high = other > mine;
        else
            high = (((unsigned)pass ^ wid) & 1u) != 0u;
        if (high)
            __builtin_amdgcn_s_setprio(1);
        else
            __builtin_amdgcn_s_setprio(0);
        if (lane == 0) __hip_atomic_store(slots + wid, mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        seen = lane < (unsigned)kSimdSlots ? __hip_atomic_load(slots + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                           : 0u;
        done = mine;
    }

    __device__ __forceinline__ void finish() {
        __builtin_amdgcn_s_setprio(0);
        if ((threadIdx.x & 63u) == 0u) __hip_atomic_store(slots + wid, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    constexpr int kXcd = 8;
    const int xcd = b % kXcd, local = b / kXcd;
    const int per = nblocks / kXcd, rem = nblocks % kXcd;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

}  // namespace gd2
