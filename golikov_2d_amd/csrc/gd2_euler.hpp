// gd2_euler.hpp -- fused Euler stage kernel for gfx950.
//
// One 64-lane wave owns a strip of 63 interior columns and marches up a run
// of rows. Lane l holds column cbase + l:
//   * y-direction: the lane keeps rows j, j+1, j+2 of its own column in
//     registers, so every y-face flux G(j+1/2) is computed once and carried
//     to the next row;
//   * x-direction: each loaded row (cbase-2 .. cbase+64, converted to
//     primitives once) goes to a 2-row LDS ring; lane l computes the x-face
//     F(c-1/2) from LDS and receives F(c+1/2) from lane l+1 (lane 63 only
//     supplies that face), so every x-face is also computed once.
// A stage reads each state value of the strip once from HBM (plus a 2-cell
// halo) and writes each output once: reconstruct -> flux -> update fused.
//
// Layouts: kSoA = native SoA conserved planes with an optional RK
// combination and wave-speed epilogue; kAoS = the reference AoS primitive
// layout with its primitive-form update (kernel.cu:1859-1874), its lagged
// face time step and frozen edge cells.
#pragma once

#include "gd2_device.hpp"

namespace gd2 {

constexpr int kTileW = 63;   // interior columns per wave
constexpr int kRowBuf = 68;  // LDS row: columns cbase-2 .. cbase+64 (67 used)

enum { kSoA = 0, kAoS = 1 };

struct EulerArgs {
    // state
    const double* Uin;     // kSoA: 4 planes
    const double* Ubase;   // kSoA: 4 planes or null (a == 0)
    double* Uout;          // kSoA: 4 planes
    double* Ustage;        // one-pass step: receives the stage-1 state U1 (may be null)
    const double2* s;      // kAoS: (rho, p)
    const double2* u;      // kAoS: (vx, vy)
    double2* s2;
    double2* u2;
    // grid
    int nx, ny;
    long long plane;
    double dx, dy, idx, idy;
    double gamma, g1, inv_g1, krit;
    int bc[4];
    // time stepping
    double a, b;
    double dt;                     // used when dt_dev == null and clk == null
    const double* dt_dev;
    // one-pass step with its clock update inside (no end time): dt =
    // clock_dt(clk, clk_slot) in every wave, clock_commit by one thread
    Clock* clk;
    int clk_slot;
    double* dt_log;
    long long dt_cap;
    unsigned long long* wmax_bits; // kSoA: wave-speed max of the output (may be null)
    unsigned long long* tmin_bits; // kAoS: face time-step min (the reference's *T)
    DeviceDiag* diag;
    // tiling
    int tiles_x, rows;
    // Kernel_TVD_old's options (kAoS, order 2): inner frozen sphere
    // (kernel.cu:3303) and axisymmetric 1/y source terms (kernel.cu:3651-3662)
    double x0, y0, r_frozen;
    int axisym;
    // pressure 1/(gamma M_inf^2) of Kernel_TVD_old's west ghost (kernel.cu:3326)
    double p_inflow;
};

// Which of Kernel_TVD_old's own rules an order-2 kAoS instance applies: none,
// the options (frozen sphere, 1/y terms), or the options and its west ghost.
enum { kRefPlain = 0, kRefOptions = 1, kRefWestInflow = 2 };

// Kernel_TVD_old's west ghost from the edge cell q (kernel.cu:3323-3337): its
// density and vy, the pressure 1/(gamma M_inf^2), and vx = -5 where the edge
// cell is subsonic with vx > -5. The Mach number is formed without
// contraction, term by term as written there.
__device__ __forceinline__ PrimE west_subsonic_rule(const EulerArgs& A, PrimE q) {
#pragma clang fp contract(off)
    const double mach = sqrt((q.u * q.u + q.v * q.v) / (A.gamma * q.p / q.r));
    if ((q.u > -5.0) && (mach < 1.0)) q.u = -5.0;
    return q;
}

struct Slope {
    double r, u, v, p;
};

template <int LAYOUT>
struct RowVal {
    PrimE q;
};

// Raw values of one cell as stored: conserved (kSoA) or (rho, p, vx, vy) (kAoS).
struct Raw {
    double a[4];
};

// Byte offset that puts a buffer access out of range: the load returns 0 and
// touches no memory. Lanes without a cell to load use it, so every load of a
// row is issued by the whole wave without a branch and the wait counts stay
// exact.
constexpr unsigned kOffNone = 0x80000000u;

// Buffer resources of one logical row (after the boundary map): one per
// conserved plane (kSoA), or s and u twice (kAoS, second double at +8 bytes).
struct RowPtr {
    __amdgpu_buffer_rsrc_t f[4];
    unsigned add[4];
};

template <int LAYOUT>
__device__ __forceinline__ RowPtr row_ptr(const EulerArgs& A, int jj) {
    RowPtr r;
    if constexpr (LAYOUT == kSoA) {
        const double* b = A.Uin + (long long)jj * A.nx;
        r.f[0] = buf_rsrc(uniform_ptr(b));
        r.f[1] = buf_rsrc(uniform_ptr(b + A.plane));
        r.f[2] = buf_rsrc(uniform_ptr(b + 2 * A.plane));
        r.f[3] = buf_rsrc(uniform_ptr(b + 3 * A.plane));
        r.add[0] = r.add[1] = r.add[2] = r.add[3] = 0u;
    } else {
        r.f[0] = r.f[1] = buf_rsrc(uniform_ptr(A.s + (long long)jj * A.nx));
        r.f[2] = r.f[3] = buf_rsrc(uniform_ptr(A.u + (long long)jj * A.nx));
        r.add[0] = r.add[2] = 0u;
        r.add[1] = r.add[3] = 8u;
    }
    return r;
}

// off: byte offset of the cell within the row (8 * column for kSoA, 16 * column for kAoS)
__device__ __forceinline__ Raw load_raw(const RowPtr& r, unsigned off) {
    Raw v;
    v.a[0] = ld_buf(r.f[0], off + r.add[0]);
    v.a[1] = ld_buf(r.f[1], off + r.add[1]);
    v.a[2] = ld_buf(r.f[2], off + r.add[2]);
    v.a[3] = ld_buf(r.f[3], off + r.add[3]);
    return v;
}

template <int LAYOUT>
__device__ __forceinline__ RowVal<LAYOUT> finish_cell(const EulerArgs& A, const Raw& w, bool fx, bool fy) {
    RowVal<LAYOUT> v;
    if constexpr (LAYOUT == kSoA) {
        v.q = cons_to_prim_e(w.a[0], w.a[1], w.a[2], w.a[3], A.g1);
    } else {
        v.q.r = w.a[0];
        v.q.p = w.a[1];
        v.q.u = w.a[2];
        v.q.v = w.a[3];
    }
    if (fx) v.q.u = -v.q.u;
    if (fy) v.q.v = -v.q.v;
    return v;
}

// Per-wave LDS: primitive rows (own cells and halo) for the x-faces and the
// conserved own cells (kSoA) waiting for their update, each a ring of two
// rows (row r in slot r & 1): 8.4 KB per wave.
struct TileLds {
    double q[2][4][kRowBuf];
    double u[2][4][kWave];
};

__device__ __forceinline__ void lds_put(double (*ring)[4][kRowBuf], int slot, int e, const PrimE& q) {
    ring[slot][0][e] = q.r;
    ring[slot][1][e] = q.u;
    ring[slot][2][e] = q.v;
    ring[slot][3][e] = q.p;
}

// GUARD = false: every cell of the face's stencil has rho, p > kPosMin, so the
// face states are positive and the zero guards can be skipped.
template <int FLUX, bool GUARD = true>
__device__ __forceinline__ FluxE face_flux(const EulerArgs& A, double rL, double unL, double utL, double pL,
                                           double rR, double unR, double utR, double pR, double& smax) {
    if constexpr (FLUX == GD2_FLUX_EXACT)
        return exact_euler(rL, unL, utL, pL, rR, unR, utR, pR, A.gamma, smax, A.diag);
    else if constexpr (FLUX == GD2_FLUX_HLL)
        return hll_euler<GUARD>(rL, unL, utL, pL, rR, unR, utR, pR, A.gamma, A.inv_g1, smax);
    else if constexpr (FLUX == GD2_FLUX_HLLC_KOROLKOV)
        return hllc_korolkov_euler<GUARD>(rL, unL, utL, pL, rR, unR, utR, pR, A.gamma, A.inv_g1, smax);
    else if constexpr (FLUX == GD2_FLUX_HLLC_2D)
        return hllc_toro_euler<GUARD>(rL, unL, utL, pL, rR, unR, utR, pR, A.gamma, A.inv_g1, smax);
    else
        return hllc_euler<GUARD>(rL, unL, utL, pL, rR, unR, utR, pR, A.gamma, A.inv_g1, smax);
}

// Cells with rho and p above this bound make every minmod face value
// positive: for q, q_- , q_+ > 0 the half slope is at most q/2 towards zero,
// so q +- s/2 >= q/2 > 0 in exact and in rounded arithmetic (rounding is
// monotone), and the products in the sound speed stay normal.
constexpr double kPosMin = 1e-150;

// minmod slope of the middle value; half-slope face values with the
// reference's rho/p <= 0 fall-back to the cell value (grad_p, kernel.cu:28,
// 3436-3544).
template <int ORDER>
__device__ __forceinline__ Slope slope3(const PrimE& a, const PrimE& b, const PrimE& c) {
    Slope s;
    if constexpr (ORDER == 2) {
        s.r = minmod(b.r - a.r, c.r - b.r);
        s.u = minmod(b.u - a.u, c.u - b.u);
        s.v = minmod(b.v - a.v, c.v - b.v);
        s.p = minmod(b.p - a.p, c.p - b.p);
    } else {
        s.r = s.u = s.v = s.p = 0.0;
    }
    return s;
}

// GUARD = false: the stencil is positive (kPosMin), the fall-back cannot fire.
template <bool GUARD = true>
__device__ __forceinline__ PrimE face_plus(const PrimE& q, const Slope& s) {
    PrimE f;
    f.r = q.r + 0.5 * s.r;
    f.u = q.u + 0.5 * s.u;
    f.v = q.v + 0.5 * s.v;
    f.p = q.p + 0.5 * s.p;
    if constexpr (GUARD) {
        if (f.r <= 0.0) f.r = q.r;
        if (f.p <= 0.0) f.p = q.p;
    }
    return f;
}

template <bool GUARD = true>
__device__ __forceinline__ PrimE face_minus(const PrimE& q, const Slope& s) {
    PrimE f;
    f.r = q.r - 0.5 * s.r;
    f.u = q.u - 0.5 * s.u;
    f.v = q.v - 0.5 * s.v;
    f.p = q.p - 0.5 * s.p;
    if constexpr (GUARD) {
        if (f.r <= 0.0) f.r = q.r;
        if (f.p <= 0.0) f.p = q.p;
    }
    return f;
}

// Flux balance of a cell per conserved field (rho, m_x, m_y, E): the sum of
// outward fluxes times face lengths (kernel.cu:1854-1857), differenced per
// direction so that equal opposite fluxes cancel exactly. The x part
// (F(c+1/2) - F(c-1/2)) dy is formed as soon as the x-faces are known; the y
// part is added with one explicit multiply-add, so every rounding is fixed by
// the source and the one-pass step reproduces the stage-by-stage one bitwise.
struct Bal {
    double r, mx, my, e;
};

__device__ __forceinline__ Bal x_balance(const FluxE& Fl, const FluxE& Fr, double dy) {
    Bal X;
    X.r = (Fr.fr - Fl.fr) * dy;
    X.mx = (Fr.fn - Fl.fn) * dy;
    X.my = (Fr.ft - Fl.ft) * dy;
    X.e = (Fr.fe - Fl.fe) * dy;
    return X;
}

// y-face fluxes are in the face frame (n = +y): fn is m_y, ft is m_x
__device__ __forceinline__ Bal full_balance(const Bal& X, const FluxE& Gm, const FluxE& Gp, double dx) {
    Bal P;
    P.r = fma(Gp.fr - Gm.fr, dx, X.r);
    P.mx = fma(Gp.ft - Gm.ft, dx, X.mx);
    P.my = fma(Gp.fn - Gm.fn, dx, X.my);
    P.e = fma(Gp.fe - Gm.fe, dx, X.e);
    return P;
}

// Stage outputs in conserved form: plain o = u - dtdV PS; with an RK base,
// o = H - (b dtdV) PS with H = a ub + b u (formed when u is produced).
__device__ __forceinline__ double out_plain(double u, double dtdV, double ps) { return fma(-dtdV, ps, u); }
__device__ __forceinline__ double rk_base(double a, double ub, double b, double u) { return fma(a, ub, b * u); }
__device__ __forceinline__ double out_base(double h, double bdt, double ps) { return fma(-bdt, ps, h); }

// Clamps of a stage output (kernel.cu:1862-1866, 1871-1874) in conserved
// form, counted when `count`. Without the wave-speed epilogue the pressure
// test needs no division: p <= 0 exactly when rho E - |m|^2 / 2 <= 0.
__device__ __forceinline__ void clamp_plain(const EulerArgs& A, double& o0, double o1, double o2, double& o3,
                                            bool count) {
    if (o0 <= 0.0) {
        if (count) atomicAdd(&A.diag->clamp_rho, 1ull);
        o0 = 0.0001;
    }
    if (o3 * o0 - 0.5 * (o1 * o1 + o2 * o2) <= 0.0) {
        if (count) atomicAdd(&A.diag->clamp_p, 1ull);
        const double inv = fast_rcp(o0);
        const double ke = 0.5 * (o1 * (o1 * inv) + o2 * (o2 * inv));
        o3 = 0.000001 / A.g1 + ke;
    }
}

// The same clamps with the output's cell wave-speed bound
// (|u| + c) / dx + (|v| + c) / dy folded into wmax.
__device__ __forceinline__ void clamp_wavespeed(const EulerArgs& A, double& o0, double o1, double o2, double& o3,
                                                bool count, double& wmax) {
    if (o0 <= 0.0) {
        if (count) atomicAdd(&A.diag->clamp_rho, 1ull);
        o0 = 0.0001;
    }
    const double inv = fast_rcp(o0);
    const double ux = o1 * inv, uy = o2 * inv;
    const double ke = 0.5 * (o1 * ux + o2 * uy);
    double p = A.g1 * (o3 - ke);
    if (p <= 0.0) {
        if (count) atomicAdd(&A.diag->clamp_p, 1ull);
        o3 = 0.000001 / A.g1 + ke;
        p = A.g1 * (o3 - ke);
    }
    const double cs = sqrt_ratio(A.gamma * p, o0);
    const double w = (fabs(ux) + cs) * A.idx + (fabs(uy) + cs) * A.idy;
    if (count) wmax = fmax(wmax, w);
}

template <int LAYOUT, int PROD>
__device__ __forceinline__ bool frozen_cell(const EulerArgs& A, int c, int j) {
    if constexpr (LAYOUT == kAoS) {
        bool f = (A.bc[0] == GD2_BC_FROZEN && c == 0) || (A.bc[1] == GD2_BC_FROZEN && c == A.nx - 1) ||
                 (A.bc[2] == GD2_BC_FROZEN && j == 0) || (A.bc[3] == GD2_BC_FROZEN && j == A.ny - 1);
        if (PROD && A.r_frozen > 0.0) {
            const double x = A.x0 + c * A.dx, y = A.y0 + j * A.dy;
            f = f || sqrt(x * x + y * y) < A.r_frozen;
        }
        return f;
    } else {
        return false;
    }
}

// One wave's strip: columns tx*63 .. tx*63+62, rows ty*rows .. +rows-1.
// BASE: the output is a*Ubase + b*(stage result) (an SSP-RK combination);
// otherwise it is the stage result itself. PROD (kAoS, order 2): Kernel_TVD_old's
// inner sphere and axisymmetric sources (kRefOptions), and its west ghost
// (kRefWestInflow).
template <int ORDER, int FLUX, int LAYOUT, bool WANT_W, bool BASE, int PROD = kRefPlain>
__device__ __forceinline__ void euler_tile(const EulerArgs& A, int tx, int ty, TileLds& lds) {
    const int lane = threadIdx.x;
    const int cbase = tx * kTileW;
    __syncthreads();  // a wave may run several tiles: the previous tile's LDS reads are done
    const int j0 = ty * A.rows;
    const int j1 = min(j0 + A.rows, A.ny);
    const int c = cbase + lane;

    bool fx, hfx;
    const int ci = map_index(c, A.nx, A.bc[0], A.bc[1], fx);
    const int hcol = lane == 0 ? cbase - 2 : (lane == 1 ? cbase - 1 : cbase + 64);
    const int hci = map_index(hcol, A.nx, A.bc[0], A.bc[1], hfx);
    const int hslot = lane == 0 ? 0 : (lane == 1 ? 1 : 66);
    const bool has_halo = lane < 3;
    // the first strip of a west-inflow grid: lanes 0, 1 hold the ghost columns
    const bool west_in = PROD == kRefWestInflow && tx == 0;
    const bool owns = (lane < kTileW) && (c < A.nx);
    constexpr unsigned kCellBytes = LAYOUT == kSoA ? 8u : 16u;
    const unsigned oc = (unsigned)ci * kCellBytes;   // own cell within a row
    const unsigned oh = has_halo ? (unsigned)hci * kCellBytes : kOffNone;  // halo cell within a row
    const unsigned oo = (unsigned)c * 8u;            // output cell within a row (kSoA)
    const unsigned ob_off = owns ? oo : kOffNone;    // RK base cell within a row

    const double dt = A.dt_dev ? *A.dt_dev : A.dt;
    const double dx = A.dx, dy = A.dy;
    const double dtdV = dt / (dx * dy);
    // lagged face time step radius: dy on x-faces and dx on y-faces as in
    // Ker_Dekard / Kernel_TVD_old (kernel.cu:1798-1801, 3549-3552)
    const double krx = A.krit * dy, kry = A.krit * dx;

    // logical row jl: raw loads of the own cell and (lanes 0-2) the halo cell
    struct Pending {
        Raw own, halo;
        bool fy;
    };
    auto issue_row = [&](int jl) -> Pending {
        Pending pr;
        const int jj = map_index(jl, A.ny, A.bc[2], A.bc[3], pr.fy);
        const RowPtr rp = row_ptr<LAYOUT>(A, jj);
        pr.own = load_raw(rp, oc);
        pr.halo = load_raw(rp, oh);
        return pr;
    };
    // convert a pending row; to_lds: also publish it (and the halo) in LDS
    // slot, with the own conserved values (kSoA). pos: whether every cell of
    // the row (and halo) has rho, p > kPosMin (wave-uniform).
    auto finish_row = [&](const Pending& pr, int slot, bool to_lds, bool& pos) -> RowVal<LAYOUT> {
        RowVal<LAYOUT> v = finish_cell<LAYOUT>(A, pr.own, fx, pr.fy);
        bool ok = v.q.r > kPosMin && v.q.p > kPosMin;
        if (to_lds) {
            lds_put(lds.q, slot, lane + 2, v.q);
            if constexpr (LAYOUT == kSoA) {
#pragma unroll
                for (int f = 0; f < 4; ++f) lds.u[slot][f][lane] = pr.own.a[f];
            }
            if (has_halo) {
                RowVal<LAYOUT> h = finish_cell<LAYOUT>(A, pr.halo, hfx, pr.fy);
                if (west_in && lane < 2) {
                    h.q = west_subsonic_rule(A, h.q);
                    h.q.p = A.p_inflow;
                }
                lds_put(lds.q, slot, hslot, h.q);
                ok = ok && h.q.r > kPosMin && h.q.p > kPosMin;
            }
        }
        pos = __all(ok) != 0;
        return v;
    };

    // prologue: rows j0-2, j0-1 (y-stencil only) and j0 (also in LDS). The
    // first y-face G(j0-1/2) is produced by the loop's own pass j = j0-1, so
    // every face goes through one instruction sequence and results do not
    // depend on where the strips start.
    const Pending pm2 = issue_row(j0 - 2), pm1 = issue_row(j0 - 1), pj0 = issue_row(j0);
    bool pa, pb, pc;
    RowVal<LAYOUT> vm2 = finish_row(pm2, 0, false, pa);
    RowVal<LAYOUT> v0 = finish_row(pm1, 0, false, pb);
    RowVal<LAYOUT> v1 = finish_row(pj0, j0 & 1, true, pc);
    // positivity of the last rows seen, newest in bit 0: rows j-1 .. j+2
    // (the whole stencil of the faces of pass j) are in bits 3 .. 0
    unsigned pos_hist = (pa ? 4u : 0u) | (pb ? 2u : 0u) | (pc ? 1u : 0u);
    Slope s0 = slope3<ORDER>(vm2.q, v0.q, v1.q);
    double smax;
    FluxE Gm;  // y-face flux G(j-1/2), frame (n = +y, t = +x)
    Gm.fr = Gm.fn = Gm.ft = Gm.fe = 0.0;
    double sGm = 0.0;

    double wmax = 0.0;
    // kAoS: the largest wave speed over the x-faces and over the y-faces of
    // the cells this lane updates; the face time step krit rad / S is
    // monotone in S, so the minimum over the faces is krit rad / max S
    // (bitwise), formed once at the end
    double sx = 0.0, sy = 0.0;
    bool any_upd = false;

    for (int j = j0 - 1; j < j1; ++j) {
        // loads first: row j+2 arrives while the x-face of row j is computed
        const Pending pn = issue_row(j + 2);

        // x-face F(c-1/2) of row j from cells c-2, c-1 (LDS), c (own), c+1
        // (LDS); row j reached LDS two iterations ago
        FluxE Fl, Fr;
        Bal X;
        double sFl = 0.0, sFr = 0.0;
        // row j (with its halo) was vetted when it was converted: bit 1 now
        const bool xpos = (pos_hist & 2u) != 0u;
        if (j >= j0) {
            const int sl = j & 1;
            const double (*ring)[4][kRowBuf] = lds.q;
            PrimE qa, qb, qd;
            qa.r = ring[sl][0][lane]; qa.u = ring[sl][1][lane]; qa.v = ring[sl][2][lane]; qa.p = ring[sl][3][lane];
            qb.r = ring[sl][0][lane + 1]; qb.u = ring[sl][1][lane + 1]; qb.v = ring[sl][2][lane + 1]; qb.p = ring[sl][3][lane + 1];
            qd.r = ring[sl][0][lane + 3]; qd.u = ring[sl][1][lane + 3]; qd.v = ring[sl][2][lane + 3]; qd.p = ring[sl][3][lane + 3];
            const PrimE& qc = v0.q;
            Slope sc = slope3<ORDER>(qb, qc, qd);
            // west inflow: the reference forms F(1/2) once per cell. The cell
            // of column 0 sees column -1 as the ghost (lane 1's qa), the cell
            // of column 1 sees it with column 0's pressure (kernel.cu:3365-3378):
            // lane 1 forms the first for lane 0, then the second as its own
            FluxE Fg;
            double sFg = 0.0;
            if constexpr (PROD == kRefWestInflow) {
                if (west_in) {
                    Slope sg = slope3<ORDER>(qa, qb, qc);
                    PrimE L = face_plus(qb, sg), R = face_minus(qc, sc);
                    Fg = face_flux<FLUX>(A, L.r, L.u, L.v, L.p, R.r, R.u, R.v, R.p, sFg);
                    if (lane == 1) qa.p = qb.p;
                }
            }
            Slope sb = slope3<ORDER>(qa, qb, qc);
            if (xpos) {
                PrimE L = face_plus<false>(qb, sb), R = face_minus<false>(qc, sc);
                Fl = face_flux<FLUX, false>(A, L.r, L.u, L.v, L.p, R.r, R.u, R.v, R.p, smax);
            } else {
                PrimE L = face_plus(qb, sb), R = face_minus(qc, sc);
                Fl = face_flux<FLUX>(A, L.r, L.u, L.v, L.p, R.r, R.u, R.v, R.p, smax);
            }
            wall_face_fix<FLUX>(Fl, c == 0 && A.bc[0] == GD2_BC_REFLECTIVE);
            sFl = (LAYOUT == kAoS) ? smax : 0.0;
            // F(c+1/2) from the next lane
            Fr.fr = from_next_lane(Fl.fr);
            Fr.fn = from_next_lane(Fl.fn);
            Fr.ft = from_next_lane(Fl.ft);
            Fr.fe = from_next_lane(Fl.fe);
            if constexpr (LAYOUT == kAoS) sFr = from_next_lane(sFl);
            if constexpr (PROD == kRefWestInflow) {
                if (west_in) {
                    const double gr = from_next_lane(Fg.fr), gn = from_next_lane(Fg.fn);
                    const double gt = from_next_lane(Fg.ft), ge = from_next_lane(Fg.fe);
                    const double gs = from_next_lane(sFg);
                    if (lane == 0) {
                        Fr.fr = gr; Fr.fn = gn; Fr.ft = gt; Fr.fe = ge;
                        sFr = gs;
                    }
                }
            }
            X = x_balance(Fl, Fr, dy);
        }
        // the update's inputs: row j's conserved values from LDS (stored when
        // the row was converted), the RK base from memory (issued by the whole
        // wave; the lead-in pass and lanes without a cell load nothing)
        double ub[4] = {0.0, 0.0, 0.0, 0.0};
        Raw uc;
        if constexpr (LAYOUT == kSoA) {
            if (j >= j0) {
#pragma unroll
                for (int f = 0; f < 4; ++f) uc.a[f] = lds.u[j & 1][f][lane];
            }
            if constexpr (BASE) {
                const unsigned off = j >= j0 ? ob_off : kOffNone;
                const double* b = A.Ubase + (long long)max(j, j0) * A.nx;
                ub[0] = ld_buf(buf_rsrc(uniform_ptr(b)), off);
                ub[1] = ld_buf(buf_rsrc(uniform_ptr(b + A.plane)), off);
                ub[2] = ld_buf(buf_rsrc(uniform_ptr(b + 2 * A.plane)), off);
                ub[3] = ld_buf(buf_rsrc(uniform_ptr(b + 3 * A.plane)), off);
            }
        }

        // row j's LDS slot is read: row j+2 may take it
        __syncthreads();
        bool pn_pos;
        RowVal<LAYOUT> v2 = finish_row(pn, (j + 2) & 1, true, pn_pos);
        pos_hist = ((pos_hist << 1) | (pn_pos ? 1u : 0u)) & 15u;
        // y-face G(j+1/2) from rows j-1 .. j+2
        Slope s1 = slope3<ORDER>(v0.q, v1.q, v2.q);
        FluxE Gp;
        double sGp;
        if (pos_hist == 15u) {
            PrimE L = face_plus<false>(v0.q, s0), R = face_minus<false>(v1.q, s1);
            Gp = face_flux<FLUX, false>(A, L.r, L.v, L.u, L.p, R.r, R.v, R.u, R.p, smax);
        } else {
            PrimE L = face_plus(v0.q, s0), R = face_minus(v1.q, s1);
            Gp = face_flux<FLUX>(A, L.r, L.v, L.u, L.p, R.r, R.v, R.u, R.p, smax);
        }
        wall_face_fix<FLUX>(Gp, j == -1 && A.bc[2] == GD2_BC_REFLECTIVE);
        sGp = (LAYOUT == kAoS) ? smax : 0.0;
        if (j < j0) {  // lead-in pass: only G(j0-1/2)
            v0 = v1;
            v1 = v2;
            s0 = s1;
            Gm = Gp;
            sGm = sGp;
            continue;
        }

        if (owns && !frozen_cell<LAYOUT, PROD>(A, c, j)) {
            const Bal P = full_balance(X, Gm, Gp, dx);
            if constexpr (LAYOUT == kSoA) {
                double o0, o1, o2, o3;
                if constexpr (BASE) {
                    const double bdt = A.b * dtdV;
                    o0 = out_base(rk_base(A.a, ub[0], A.b, uc.a[0]), bdt, P.r);
                    o1 = out_base(rk_base(A.a, ub[1], A.b, uc.a[1]), bdt, P.mx);
                    o2 = out_base(rk_base(A.a, ub[2], A.b, uc.a[2]), bdt, P.my);
                    o3 = out_base(rk_base(A.a, ub[3], A.b, uc.a[3]), bdt, P.e);
                } else {
                    o0 = out_plain(uc.a[0], dtdV, P.r);
                    o1 = out_plain(uc.a[1], dtdV, P.mx);
                    o2 = out_plain(uc.a[2], dtdV, P.my);
                    o3 = out_plain(uc.a[3], dtdV, P.e);
                }
                if constexpr (WANT_W)
                    clamp_wavespeed(A, o0, o1, o2, o3, true, wmax);
                else
                    clamp_plain(A, o0, o1, o2, o3, true);
                double* ob = A.Uout + (long long)j * A.nx;
                st_buf(buf_rsrc(uniform_ptr(ob)), oo, o0);
                st_buf(buf_rsrc(uniform_ptr(ob + A.plane)), oo, o1);
                st_buf(buf_rsrc(uniform_ptr(ob + 2 * A.plane)), oo, o2);
                st_buf(buf_rsrc(uniform_ptr(ob + 3 * A.plane)), oo, o3);
            } else {
                // primitive-form update of Ker_Dekard (kernel.cu:1859-1874);
                // Kernel_TVD_old's 1/y terms with the axisymmetric option
                // (kernel.cu:3651-3662). Divisions by gamma-1, y and the new
                // density as products with reciprocals (rounding-level).
                const long long k = (long long)j * A.nx + c;
                const PrimE& q = v0.q;
                const double gm1 = A.gamma - 1.0;
                const double Tdv = dt / (dx * dy);
                const double PSr = P.r, PUx = P.mx, PUy = P.my, PSe = P.e;
                double r2, ux, uy, p2;
                if (PROD && ORDER == 2 && A.axisym) {
                    const double dty = dt / (A.y0 + j * dy);
                    r2 = q.r - Tdv * PSr - dty * q.r * q.v;
                    if (r2 <= 0.0) {
                        atomicAdd(&A.diag->clamp_rho, 1ull);
                        r2 = 0.0001;
                    }
                    const double ir2 = 1.0 / r2;
                    ux = (q.r * q.u - Tdv * PUx - dty * q.r * q.v * q.u) * ir2;
                    uy = (q.r * q.v - Tdv * PUy - dty * q.r * q.v * q.v) * ir2;
                    p2 = (((q.p * A.inv_g1 + q.r * (q.u * q.u + q.v * q.v) * 0.5) - Tdv * PSe -
                           dty * q.v * (A.gamma * q.p * A.inv_g1 + q.r * (q.u * q.u + q.v * q.v) * 0.5)) -
                          0.5 * r2 * (ux * ux + uy * uy)) * gm1;
                } else {
                    r2 = q.r - Tdv * PSr;
                    if (r2 <= 0.0) {
                        atomicAdd(&A.diag->clamp_rho, 1ull);
                        r2 = 0.0001;
                    }
                    const double ir2 = 1.0 / r2;
                    ux = (q.r * q.u - Tdv * PUx) * ir2;
                    uy = (q.r * q.v - Tdv * PUy) * ir2;
                    p2 = (((q.p * A.inv_g1 + q.r * (q.u * q.u + q.v * q.v) * 0.5) - Tdv * PSe) -
                          0.5 * r2 * (ux * ux + uy * uy)) * gm1;
                }
                if (p2 <= 0.0) {
                    atomicAdd(&A.diag->clamp_p, 1ull);
                    p2 = 0.000001;
                }
                A.s2[k] = make_double2(r2, p2);
                A.u2[k] = make_double2(ux, uy);
                // per-cell tmin starts at 1000 (kernel.cu:1748): applied at the end
                sx = fmax(sx, fmax(sFr, sFl));
                sy = fmax(sy, fmax(sGm, sGp));
                any_upd = true;
            }
        } else if constexpr (LAYOUT == kAoS) {
            if (owns) {  // frozen edge cell: carried over unchanged
                const long long k = (long long)j * A.nx + c;
                A.s2[k] = A.s[k];
                A.u2[k] = A.u[k];
            }
        }

        // advance the y-window
        v0 = v1;
        v1 = v2;
        s0 = s1;
        Gm = Gp;
        sGm = sGp;
    }

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

template <int ORDER, int FLUX, int LAYOUT, bool WANT_W, bool BASE, int PROD = kRefPlain>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu((FLUX == GD2_FLUX_HLLC || FLUX == GD2_FLUX_HLLC_KOROLKOV || FLUX == GD2_FLUX_HLLC_2D) ? (LAYOUT == kAoS && ORDER == 2 ? 4 : 3) : 1))) void euler_tile_kernel(const EulerArgs A) {
    __shared__ TileLds lds;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    euler_tile<ORDER, FLUX, LAYOUT, WANT_W, BASE, PROD>(A, tile % A.tiles_x, tile / A.tiles_x, lds);
}

// Cell-centred wave-speed bound max((|u|+c)/dx + (|v|+c)/dy) of a SoA
// Euler state, folded into *wmax_bits.
static __global__ __launch_bounds__(256) void euler_wavespeed_kernel(const double* U, long long n, long long plane,
                                                              double gamma, double g1, double idx, double idy,
                                                              unsigned long long* wmax_bits) {
    double w = 0.0;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        PrimE q = cons_to_prim_e(U[k], U[plane + k], U[2 * plane + k], U[3 * plane + k], g1);
        const double cs = sqrt(gamma * q.p / q.r);
        w = fmax(w, (fabs(q.u) + cs) * idx + (fabs(q.v) + cs) * idy);
    }
    // wave max, then block max through LDS, then one atomic per block
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
