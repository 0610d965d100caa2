// gd2_adjoint.hpp -- discrete adjoint (vector-Jacobian product) of the Euler
// HLLC + minmod-MUSCL stage operator on gfx950.
//
// For a stage operator L(U) = -(1/dV) sum_faces (outward flux x face length)
// the kernel computes, per cell,
//     W = a*A + b*B + c * J_L(U)^T V
// with V a cotangent field (4 SoA planes). J_L^T V is formed face by face:
// the face cotangent phi = (V_R - V_L)/h goes through the reverse-mode HLLC
// (hllc_vjp) to cotangents on the two reconstructed face states, which each
// cell spreads over its 3-cell minmod stencil with the branch taken in the
// forward pass (subgradient 1/2-1/2 on ties, 0 at the zero branch). Cells
// gather their contributions without atomics: x-neighbours by cross-lane
// shuffles, y-neighbours by a register march that finishes row j one row
// later. Cotangents on ghost cells go to a small side buffer and are folded
// into the cells they copy (transmissive / reflective / periodic) by
// adj_fold_kernel.
#pragma once

#include <type_traits>

#include "gd2_euler.hpp"

namespace gd2 {

constexpr int kAdjW = 61;    // interior columns per wave (lanes 1..61)
constexpr int kAdjBuf = 68;  // LDS row: columns cbase-3 .. cbase+63

struct FaceE {
    double r, un, ut, p;
};

// Reverse-mode HLLC (same forward formulas as hllc_euler): given the flux
// cotangent Fb (mass, normal, tangential, energy) return the cotangents of
// the left and right face states.
template <bool GUARD = true>
__device__ __forceinline__ void hllc_vjp(const FaceE& L, const FaceE& R, double gamma, double inv_g1, const FluxE& Fb,
                                         FaceE& Lb, FaceE& Rb) {
    // forward
    const SoundSpeed sL = sound_speed<GUARD>(gamma * L.p, L.r), sR = sound_speed<GUARD>(gamma * R.p, R.r);
    const double cL = sL.c, cR = sR.c;
    const double aL = L.un - cL, aR = R.un - cR;
    const double bL = L.un + cL, bR = R.un + cR;
    const double SL = fmin(aL, aR), SR = fmax(bL, bR);
    const bool slL = aL <= aR, srL = bL >= bR;
    const double suL = SL - L.un, suR = SR - R.un;
    const double mL = suL * L.r, mR = suR * R.r;
    const double smR = mR * R.un;
    const double smL = mL * L.un;
    const double D = mR - mL;
    const double iD = fast_rcp(D);
    const double Nn = smR - R.p + L.p - smL;
    const double SM = Nn * iD;
    const double ptz = 0.5 * ((R.p + mR * (SM - R.un)) + (L.p + mL * (SM - L.un)));
    const bool left = SM >= 0.0;
    const double S = left ? SL : SR;
    const double r = left ? L.r : R.r;
    const double un = left ? L.un : R.un;
    const double ut = left ? L.ut : R.ut;
    const double p = left ? L.p : R.p;
    const double su = S - un;
    const bool super = (SL > 0.0) || (SR < 0.0);
    const double ke = 0.5 * (un * un + ut * ut);
    const double e = p * inv_g1 + r * ke;
    const double den = super ? 1.0 : (S - SM);
    const double inv = fast_rcp(den);
    const double sum = su * inv;
    const double rz = r * sum;
    const double t = ptz * SM - p * un;
    const double ez = e * sum + t * inv;
    const double Se = super ? 0.0 : S;
    const double run = r * un;

    // reverse
    const double b_Se = Fb.fr * (rz - r) + Fb.fn * (SM * rz - run) + Fb.ft * (ut * rz - r * ut) + Fb.fe * (ez - e);
    double b_run = Fb.fr + Fb.fn * (un - Se) + Fb.ft * ut;
    const double b_rz = Se * (Fb.fr + Fb.fn * SM + Fb.ft * ut);
    double b_r = -Se * (Fb.fr + Fb.ft * ut);
    double b_un = Fb.fn * run + Fb.fe * (e + p);
    double b_p = Fb.fn + Fb.fe * un;
    double b_SM = Fb.fn * Se * rz;
    double b_ut = Fb.ft * (run + Se * (rz - r));
    const double b_ez = Fb.fe * Se;
    double b_e = Fb.fe * (un - Se);
    b_r += b_run * un;
    b_un += b_run * r;
    double b_S = super ? 0.0 : b_Se;
    b_e += b_ez * sum;
    double b_sum = b_ez * e;
    const double b_t = b_ez * inv;
    double b_inv = b_ez * t;
    const double b_ptz = b_t * SM;
    b_SM += b_t * ptz;
    b_p -= b_t * un;
    b_un -= b_t * p;
    b_r += b_rz * sum;
    b_sum += b_rz * r;
    const double b_su = b_sum * inv;
    b_inv += b_sum * su;
    const double b_den = -inv * inv * b_inv;
    if (!super) {
        b_S += b_den;
        b_SM -= b_den;
    }
    b_p += b_e * inv_g1;
    b_r += b_e * ke;
    b_un += b_e * r * un;
    b_ut += b_e * r * ut;

    // route the selected side's cotangents with 0/1 weights (exact: x*1, x*0)
    const double wl = left ? 1.0 : 0.0, wr = left ? 0.0 : 1.0;
    FaceE bl = {b_r * wl, b_un * wl, b_ut * wl, b_p * wl};
    FaceE br = {b_r * wr, b_un * wr, b_ut * wr, b_p * wr};
    double b_SL = b_S * wl, b_SR = b_S * wr, b_suL = b_su * wl, b_suR = b_su * wr;
    const double hp = 0.5 * b_ptz;
    br.p += hp;
    double b_mR = hp * (SM - R.un);
    b_SM += hp * (mR + mL);
    br.un -= hp * mR;
    bl.p += hp;
    double b_mL = hp * (SM - L.un);
    bl.un -= hp * mL;
    const double b_N = b_SM * iD;
    const double b_iD = b_SM * Nn;
    br.p -= b_N;
    bl.p += b_N;
    const double b_D = -iD * iD * b_iD;
    b_mR += b_D;
    b_mL -= b_D;
    b_mR += b_N * R.un;
    br.un += b_N * mR;
    b_mL -= b_N * L.un;
    bl.un -= b_N * mL;
    b_suL += b_mL * L.r;
    bl.r += b_mL * suL;
    b_suR += b_mR * R.r;
    br.r += b_mR * suR;
    b_SL += b_suL;
    bl.un -= b_suL;
    b_SR += b_suR;
    br.un -= b_suR;
    // SL = min(uL - cL, uR - cR), SR = max(uL + cL, uR + cR): the winning side's
    // share, split exactly with 0/1 weights
    const double t1 = b_SL * (slL ? 1.0 : 0.0), t2 = b_SR * (srL ? 1.0 : 0.0);
    const double o1 = b_SL - t1, o2 = b_SR - t2;
    bl.un += t1 + t2;
    br.un += o1 + o2;
    const double b_cL = t2 - t1, b_cR = o2 - o1;
    // c = sqrt(gamma p / r): dc/dp = gamma h, dc/dr = -c^2 h (h = 1 / (2 sqrt(gamma p r)))
    bl.p += b_cL * (gamma * sL.h);
    bl.r -= b_cL * (cL * cL * sL.h);
    br.p += b_cR * (gamma * sR.h);
    br.r -= b_cR * (cR * cR * sR.h);
    Lb = bl;
    Rb = br;
}

// minmod branch weights of the middle cell's slope: the slope is
// alpha*(q_c - q_{c-1}) + beta*(q_{c+1} - q_c) on the branch taken
// (alpha = 1: first argument, beta = 1: second, 1/2-1/2 on a tie, 0-0 when the
// arguments' signs differ or one is zero).
struct MMW {
    double a, b;
};

__device__ __forceinline__ MMW mm_weights(double dm, double dp) {
    MMW w;
    const double am = fabs(dm), ap = fabs(dp);
    const bool zero = !(dm * dp > 0.0);
    w.a = zero ? 0.0 : (am < ap ? 1.0 : (am > ap ? 0.0 : 0.5));
    w.b = zero ? 0.0 : (ap < am ? 1.0 : (ap > am ? 0.0 : 0.5));
    return w;
}

// One cell's share of the reconstruction adjoint along one direction and one
// field: given the cotangents of its plus-face value (qp = q + mm/2) and its
// minus-face value (qm = q - mm/2), with fp / fm set where the rho/p <= 0
// fall-back replaced the face value by the cell value, return the parts that
// belong to the cell itself, to the previous cell and to the next cell.
__device__ __forceinline__ void spread(double dqp, double dqm, const MMW& w, bool fp, bool fm, double& own,
                                       double& to_prev, double& to_next) {
    const double g = 0.5 * ((fp ? 0.0 : dqp) - (fm ? 0.0 : dqm));
    own = dqp + dqm + (w.a - w.b) * g;
    to_prev = -w.a * g;
    to_next = w.b * g;
}

struct AdjArgs {
    const double* U;    // state the stage operator is linearised at
    const double* V;    // cotangent
    const double* A;    // optional
    const double* B;    // optional
    double* W;          // output
    double a, b, c;     // W = a*A + b*B + c*J^T V
    double* ghost;      // side buffer: [4 sides][2 layers][len][4]
    int nx, ny;
    long long plane;
    double dx, dy, idx, idy;
    double gamma, g1, inv_g1;
    int bc[4];
    int tiles_x, rows;
    int gstride;        // max(nx, ny)
    // 1: the kernel adds the ghost-cell cotangents of transmissive and
    // reflective sides into the cells they copy itself (every such cell lies
    // in the same wave or strip as its ghosts); 0: they go to the side buffer
    // for adj_fold_kernel
    int fold_in_kernel;
};

struct Q4 {
    double r, u, v, p;
};

__device__ __forceinline__ Q4 q4_zero() {
    Q4 z = {0.0, 0.0, 0.0, 0.0};
    return z;
}

__device__ __forceinline__ Q4 q4_add(const Q4& a, const Q4& b) {
    Q4 s = {a.r + b.r, a.u + b.u, a.v + b.v, a.p + b.p};
    return s;
}

// Q4 of lane + 1 (lane 63 receives 0) and of lane - 1 (lane 0 receives 0).
__device__ __forceinline__ Q4 q4_from_next(const Q4& x) {
    Q4 y = {from_next_lane(x.r), from_next_lane(x.u), from_next_lane(x.v), from_next_lane(x.p)};
    return y;
}

__device__ __forceinline__ Q4 q4_from_prev(const Q4& x) {
    Q4 y = {from_prev_lane(x.r), from_prev_lane(x.u), from_prev_lane(x.v), from_prev_lane(x.p)};
    return y;
}

// Cell-level reconstruction adjoint for one direction: slopes from
// (qm1, q, qp1), face-value fall-backs from the forward formulas.
template <int ORDER, bool GUARD = true>
__device__ __forceinline__ void spread_cell(const PrimE& qm1, const PrimE& q, const PrimE& qp1, const Q4& dqp,
                                            const Q4& dqm, Q4& own, Q4& to_prev, Q4& to_next) {
    if constexpr (ORDER == 1) {
        own = q4_add(dqp, dqm);
        to_prev = q4_zero();
        to_next = q4_zero();
    } else {
        const MMW wr = mm_weights(q.r - qm1.r, qp1.r - q.r);
        const MMW wu = mm_weights(q.u - qm1.u, qp1.u - q.u);
        const MMW wv = mm_weights(q.v - qm1.v, qp1.v - q.v);
        const MMW wp = mm_weights(q.p - qm1.p, qp1.p - q.p);
        bool fpr = false, fmr = false, fpp = false, fmp = false;
        if constexpr (GUARD) {
            const Slope s = slope3<2>(qm1, q, qp1);
            fpr = !(q.r + 0.5 * s.r > 0.0);
            fmr = !(q.r - 0.5 * s.r > 0.0);
            fpp = !(q.p + 0.5 * s.p > 0.0);
            fmp = !(q.p - 0.5 * s.p > 0.0);
        }
        spread(dqp.r, dqm.r, wr, fpr, fmr, own.r, to_prev.r, to_next.r);
        spread(dqp.u, dqm.u, wu, false, false, own.u, to_prev.u, to_next.u);
        spread(dqp.v, dqm.v, wv, false, false, own.v, to_prev.v, to_next.v);
        spread(dqp.p, dqm.p, wp, fpp, fmp, own.p, to_prev.p, to_next.p);
    }
}

// d(prim)/d(cons) transposed at a cell with density r and velocity (u, v):
// prim = (r, mx/r, my/r, g1 (E - (mx^2 + my^2)/(2 r))).
__device__ __forceinline__ void cons_vjp(double r, double u, double v, double g1, const Q4& dq, double out[4]) {
    const double inv = fast_rcp(r);
    out[3] = g1 * dq.p;
    out[1] = dq.u * inv - g1 * u * dq.p;
    out[2] = dq.v * inv - g1 * v * dq.p;
    out[0] = dq.r - (u * dq.u + v * dq.v) * inv + g1 * 0.5 * (u * u + v * v) * dq.p;
}

__device__ __forceinline__ void lds_put8(double (*ring)[8][kAdjBuf], int slot, int e, const PrimE& q, const double* V) {
    ring[slot][0][e] = q.r;
    ring[slot][1][e] = q.u;
    ring[slot][2][e] = q.v;
    ring[slot][3][e] = q.p;
    ring[slot][4][e] = V[0];
    ring[slot][5][e] = V[1];
    ring[slot][6][e] = V[2];
    ring[slot][7][e] = V[3];
}

__device__ __forceinline__ PrimE lds_q(double (*ring)[8][kAdjBuf], int slot, int e) {
    PrimE q;
    q.r = ring[slot][0][e];
    q.u = ring[slot][1][e];
    q.v = ring[slot][2][e];
    q.p = ring[slot][3][e];
    return q;
}

// ghost side buffer index: side 0 west, 1 east, 2 south, 3 north; layer 0 is
// the ghost next to the boundary, layer 1 the outer one
__device__ __forceinline__ double* ghost_at(const AdjArgs& A, int side, int layer, int idx) {
    return A.ghost + (((long long)(side * 2 + layer) * A.gstride + idx) * 4);
}

__device__ __forceinline__ void ghost_store(const AdjArgs& A, int side, int layer, int idx, const Q4& d) {
    double* g = ghost_at(A, side, layer, idx);
    g[0] = d.r;
    g[1] = d.u;
    g[2] = d.v;
    g[3] = d.p;
}

// x of lane `src` (wave-uniform) in every lane.
__device__ __forceinline__ double lane_value(double x, int src) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ Q4 q4_lane(const Q4& x, int src) {
    Q4 y = {lane_value(x.r, src), lane_value(x.u, src), lane_value(x.v, src), lane_value(x.p, src)};
    return y;
}

__device__ __forceinline__ Q4 q4_flip_u(const Q4& x, bool f) {
    Q4 y = {x.r, f ? -x.u : x.u, x.v, x.p};
    return y;
}

template <int ORDER>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void euler_vjp_kernel(const AdjArgs A) {
    __shared__ double ring[3][8][kAdjBuf];
    __shared__ double gyb[2][4][kWave];  // ghost-row cotangents (layer 0 next to the edge)
    const int lane = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % A.tiles_x, ty = tile / A.tiles_x;
    const int cbase = tx * kAdjW;
    const int j0 = ty * A.rows;
    const int j1 = min(j0 + A.rows, A.ny);
    const int c = cbase - 1 + lane;
    const int nx = A.nx, ny = A.ny;
    const long long P = A.plane;

    bool fx, hfx;
    const int ci = map_index(c, nx, A.bc[0], A.bc[1], fx);
    const int hcol = lane == 0 ? cbase - 3 : (lane == 1 ? cbase - 2 : cbase + 63);
    const int hci = map_index(hcol, nx, A.bc[0], A.bc[1], hfx);
    const int hslot = lane == 0 ? 0 : (lane == 1 ? 1 : 66);
    const bool has_halo = lane < 3;
    const bool col_real = c >= 0 && c < nx;
    const bool xface_on = c >= 0 && c <= nx;  // face c-1/2 belongs to the forward map
    const bool interior_lane = lane >= 1 && lane <= kAdjW;

    // the strip's output rows include the ghost rows on the domain's edges
    const int ro0 = (j0 == 0) ? -2 : j0;
    const int ro1 = (j1 == ny) ? ny + 2 : j1;

    // in-kernel folding: the cells each ghost layer copies (layer 0 is next
    // to the edge) and whether the copy mirrors the normal velocity
    int tgt_w[2], tgt_e[2], tgt_lo[2], tgt_hi[2];
    bool flip_w[2], flip_e[2], flip_lo[2], flip_hi[2];
    for (int layer = 0; layer < 2; ++layer) {
        tgt_w[layer] = map_index(-1 - layer, nx, A.bc[0], A.bc[1], flip_w[layer]);
        tgt_e[layer] = map_index(nx + layer, nx, A.bc[0], A.bc[1], flip_e[layer]);
        tgt_lo[layer] = map_index(-1 - layer, ny, A.bc[2], A.bc[3], flip_lo[layer]);
        tgt_hi[layer] = map_index(ny + layer, ny, A.bc[2], A.bc[3], flip_hi[layer]);
    }
    const bool fold_lo = A.fold_in_kernel && j0 == 0;
    const bool fold_hi = A.fold_in_kernel && j1 == ny;

    const unsigned oc = (unsigned)ci * 8u;
    const unsigned oh = has_halo ? (unsigned)hci * 8u : kOffNone;
    // ghost cells have no output in the forward map: their cotangent is 0
    // (an out-of-range offset loads 0)
    const unsigned ocv = col_real ? oc : kOffNone;
    const unsigned ohv = (has_halo && hcol >= 0 && hcol < nx) ? oh : kOffNone;

    // raw U and V values of one logical row (own cell and, lanes 0-2, the halo
    // cell); issued by the whole wave without branches
    struct RawRow {
        double u[4], uh[4], v[4], vh[4];
        bool fy;
    };
    auto issue = [&](int jl) -> RawRow {
        RawRow w;
        const int jj = map_index(jl, ny, A.bc[2], A.bc[3], w.fy);
        const bool row_real = jl >= 0 && jl < ny;
        const unsigned ov = row_real ? ocv : kOffNone, ohw = row_real ? ohv : kOffNone;
        const long long rowk = (long long)jj * nx;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const __amdgpu_buffer_rsrc_t ru = buf_rsrc(uniform_ptr(A.U + f * P + rowk));
            const __amdgpu_buffer_rsrc_t rv = buf_rsrc(uniform_ptr(A.V + f * P + rowk));
            w.u[f] = ld_buf(ru, oc);
            w.v[f] = ld_buf(rv, ov);
            w.uh[f] = ld_buf(ru, oh);
            w.vh[f] = ld_buf(rv, ohw);
        }
        return w;
    };
    // convert a raw row; to_lds: publish it (and the halo) in LDS slot.
    // pos: every cell of the row (and halo) has rho, p > kPosMin (wave-uniform)
    auto finish = [&](const RawRow& w, int slot, bool to_lds, double Vv[4], bool& pos) -> PrimE {
        PrimE q = cons_to_prim_e(w.u[0], w.u[1], w.u[2], w.u[3], A.g1);
        if (fx) q.u = -q.u;
        if (w.fy) q.v = -q.v;
        bool ok = q.r > kPosMin && q.p > kPosMin;
#pragma unroll
        for (int f = 0; f < 4; ++f) Vv[f] = w.v[f];
        if (to_lds) {
            lds_put8(ring, slot, lane + 2, q, Vv);
            if (has_halo) {
                PrimE h = cons_to_prim_e(w.uh[0], w.uh[1], w.uh[2], w.uh[3], A.g1);
                if (hfx) h.u = -h.u;
                if (w.fy) h.v = -h.v;
                ok = ok && h.r > kPosMin && h.p > kPosMin;
                lds_put8(ring, slot, hslot, h, w.vh);
            }
        }
        pos = __all(ok) != 0;
        return q;
    };

    // rows ro0-3 .. ro0-1: prims; ro0-2, ro0-1 cotangents (all loads issued
    // before the first conversion)
    double Vd[4];
    bool pa, pb, pc;
    PrimE qa, qb, qc, qd;
    double Vb[4], Vc[4];
    {
        const RawRow ra = issue(ro0 - 3), rb = issue(ro0 - 2), rc = issue(ro0 - 1);
        qa = finish(ra, 0, false, Vd, pa);  // q_{j-1} at j = ro0-2
        qb = finish(rb, 0, false, Vd, pb);  // q_j
        Vb[0] = Vd[0]; Vb[1] = Vd[1]; Vb[2] = Vd[2]; Vb[3] = Vd[3];
        qc = finish(rc, ((ro0 - 1) % 3 + 3) % 3, true, Vd, pc);  // q_{j+1}
        Vc[0] = Vd[0]; Vc[1] = Vd[1]; Vc[2] = Vd[2]; Vc[3] = Vd[3];
    }
    // positivity of the rows seen, newest in bit 0; in pass j bits 3..0 are rows j-1 .. j+2
    unsigned pos_hist = (pa ? 4u : 0u) | (pb ? 2u : 0u) | (pc ? 1u : 0u);

    Q4 dqm_j = q4_zero();     // cotangent of row j's minus (south) face value, from G(j-1/2)
    Q4 acc_prev = q4_zero();  // row j-1: everything but the spill from row j
    Q4 rsp_prev = q4_zero();  // spill of row j-1 to row j

    // two waves share a SIMD: they trade the higher issue priority every pass
    // (odd/even hardware wave slot, HW_ID bit 0) instead of letting the older
    // one finish first and leave the other to run alone
    const unsigned wslot = __builtin_amdgcn_s_getreg(4) & 1u;
    for (int j = ro0 - 2; j <= ro1; ++j) {
        if (((unsigned)j ^ wslot) & 1u)
            __builtin_amdgcn_s_setprio(1);
        else
            __builtin_amdgcn_s_setprio(0);
        __syncthreads();  // slot (j+2) % 3 (row j-1) was read by the previous pass
        const RawRow nr = issue(j + 2);
        // row j's x gather (real rows only), while row j+2's loads are in
        // flight; row j (with halo) is bit 1 before row j+2 is added
        Q4 x_j = q4_zero();
        const bool do_x = j >= j0 && j < j1;
        auto xpass = [&](auto guard) {
            constexpr bool G = decltype(guard)::value;
            const int sl = ((j % 3) + 3) % 3;
            // x-face F(c-1/2) from cells c-2, c-1 (LDS), c (own), c+1 (LDS)
            const PrimE q2 = lds_q(ring, sl, lane), q1 = lds_q(ring, sl, lane + 1), q3 = lds_q(ring, sl, lane + 3);
            const Slope sb = slope3<ORDER>(q2, q1, qb);
            const Slope sc = slope3<ORDER>(q1, qb, q3);
            const PrimE L = face_plus<G>(q1, sb), R = face_minus<G>(qb, sc);
            const double Vw0 = ring[sl][4][lane + 1], Vw1 = ring[sl][5][lane + 1];
            const double Vw2 = ring[sl][6][lane + 1], Vw3 = ring[sl][7][lane + 1];
            FluxE Fb;
            Fb.fr = xface_on ? (Vb[0] - Vw0) * A.idx : 0.0;
            Fb.fn = xface_on ? (Vb[1] - Vw1) * A.idx : 0.0;
            Fb.ft = xface_on ? (Vb[2] - Vw2) * A.idx : 0.0;
            Fb.fe = xface_on ? (Vb[3] - Vw3) * A.idx : 0.0;
            FaceE fl = {L.r, L.u, L.v, L.p}, fr = {R.r, R.u, R.v, R.p}, bl, br;
            hllc_vjp<G>(fl, fr, A.gamma, A.inv_g1, Fb, bl, br);
            const Q4 dqL = {bl.r, bl.un, bl.ut, bl.p};  // cotangent of q_plus of cell c-1
            const Q4 dqR = {br.r, br.un, br.ut, br.p};  // cotangent of q_minus of cell c
            // this lane's cell c: dqp from face c+1/2 (next lane's dqL), dqm = dqR
            const Q4 dqp_c = q4_from_next(dqL);
            Q4 own_c, lsp_c, rsp_c;
            spread_cell<ORDER, G>(q1, qb, q3, dqp_c, dqR, own_c, lsp_c, rsp_c);
            const Q4 from_next = q4_from_next(lsp_c);
            const Q4 from_prev = q4_from_prev(rsp_c);
            x_j = q4_add(q4_add(own_c, from_next), from_prev);
            if (A.fold_in_kernel) {
                // ghost columns -1, -2 (lane 0: its own part and its spill to
                // column -2) and nx, nx+1 go to the columns they copy, with u
                // mirrored when reflective
                if (tx == 0) {
                    const Q4 g1 = q4_lane(x_j, 0), g2 = q4_lane(lsp_c, 0);
                    if (c == tgt_w[0]) x_j = q4_add(x_j, q4_flip_u(g1, flip_w[0]));
                    if (c == tgt_w[1]) x_j = q4_add(x_j, q4_flip_u(g2, flip_w[1]));
                }
                if (tx == A.tiles_x - 1) {
                    const int le = nx - cbase + 1;  // lane of column nx
                    const Q4 g1 = q4_lane(x_j, le), g2 = q4_lane(x_j, le + 1);
                    if (c == tgt_e[0]) x_j = q4_add(x_j, q4_flip_u(g1, flip_e[0]));
                    if (c == tgt_e[1]) x_j = q4_add(x_j, q4_flip_u(g2, flip_e[1]));
                }
            } else {
                // ghost columns of the domain edges (east ones from the last tile only)
                if (c == -1) {
                    ghost_store(A, 0, 0, j, x_j);
                    ghost_store(A, 0, 1, j, lsp_c);
                }
                if (tx == A.tiles_x - 1) {
                    if (c == nx) ghost_store(A, 1, 0, j, x_j);
                    if (c == nx + 1) ghost_store(A, 1, 1, j, x_j);
                }
            }
        };
        if (do_x) {
            if (pos_hist & 2u)
                xpass(std::false_type{});
            else
                xpass(std::true_type{});
        }

        // the affine terms of row j-1 (finished at the end of this pass),
        // issued before the y-face (in flight during its reversal); an
        // offset that reads nothing where this lane stores nothing or the
        // term is absent
        double Aw[4], Bw[4];
        {
            const int r = j - 1;
            const bool row_out = r >= ro0 && r < ro1 && r >= 0 && r < ny;  // wave-uniform
            const bool st = row_out && col_real && interior_lane;
            const long long rowk = (long long)(row_out ? r : 0) * nx;
            const unsigned oa = (st && A.A) ? (unsigned)c * 8u : kOffNone;
            const unsigned ob = (st && A.B) ? (unsigned)c * 8u : kOffNone;
            const double* Ab = A.A ? A.A : A.U;
            const double* Bb = A.B ? A.B : A.U;
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                Aw[f] = ld_buf(buf_rsrc(uniform_ptr(Ab + f * P + rowk)), oa);
                Bw[f] = ld_buf(buf_rsrc(uniform_ptr(Bb + f * P + rowk)), ob);
            }
        }
        bool pd;
        qd = finish(nr, ((j + 2) % 3 + 3) % 3, true, Vd, pd);  // q_{j+2}
        const double Vn[4] = {Vd[0], Vd[1], Vd[2], Vd[3]};
        pos_hist = ((pos_hist << 1) | (pd ? 1u : 0u)) & 15u;

        // y-face G(j+1/2) between rows j and j+1 (forward face iff -1 <= j <= ny-1,
        // real column), then row j's y spread
        Q4 dqp_j, dqm_n, own_j, lsp_j, rsp_j;
        auto ypass = [&](auto guard) {
            constexpr bool G = decltype(guard)::value;
            const Slope sj = slope3<ORDER>(qa, qb, qc);
            const Slope sn = slope3<ORDER>(qb, qc, qd);
            const PrimE L = face_plus<G>(qb, sj), R = face_minus<G>(qc, sn);
            const bool on = col_real && j >= -1 && j <= ny - 1;
            FluxE Fb;  // face cotangent (V_top - V_bottom)/dy in the y-frame (n = y, t = x)
            Fb.fr = on ? (Vc[0] - Vb[0]) * A.idy : 0.0;
            Fb.ft = on ? (Vc[1] - Vb[1]) * A.idy : 0.0;
            Fb.fn = on ? (Vc[2] - Vb[2]) * A.idy : 0.0;
            Fb.fe = on ? (Vc[3] - Vb[3]) * A.idy : 0.0;
            FaceE fl = {L.r, L.v, L.u, L.p}, fr = {R.r, R.v, R.u, R.p}, bl, br;
            hllc_vjp<G>(fl, fr, A.gamma, A.inv_g1, Fb, bl, br);
            dqp_j.r = bl.r; dqp_j.u = bl.ut; dqp_j.v = bl.un; dqp_j.p = bl.p;
            dqm_n.r = br.r; dqm_n.u = br.ut; dqm_n.v = br.un; dqm_n.p = br.p;
            spread_cell<ORDER, G>(qa, qb, qc, dqp_j, dqm_j, own_j, lsp_j, rsp_j);
        };
        if (pos_hist == 15u)
            ypass(std::false_type{});
        else
            ypass(std::true_type{});

        // finish row r = j-1: its accumulated parts + the spill from row j
        const int r = j - 1;
        if (r >= ro0 && r < ro1) {
            Q4 tot = q4_add(acc_prev, lsp_j);
            if (r >= 0 && r < ny) {
                if (col_real && interior_lane) {
                    if (fold_lo && r <= 1) {
                        // rows -1, -2 copy rows map(-1), map(-2) (mirrored v when reflective)
                        for (int layer = 0; layer < 2; ++layer) {
                            if (tgt_lo[layer] != r) continue;
                            Q4 e = {gyb[layer][0][lane], gyb[layer][1][lane], gyb[layer][2][lane],
                                    gyb[layer][3][lane]};
                            if (flip_lo[layer]) e.v = -e.v;
                            tot = q4_add(tot, e);
                        }
                    }
                    double du[4];
                    cons_vjp(qa.r, qa.u, qa.v, A.g1, tot, du);  // qa = prims of row j-1
                    const long long rowk = (long long)r * nx;
                    const unsigned o = (unsigned)c * 8u;
#pragma unroll
                    for (int f = 0; f < 4; ++f) {
                        double w = A.c * du[f];
                        if (A.A) w += A.a * Aw[f];
                        if (A.B) w += A.b * Bw[f];
                        st_buf(buf_rsrc(uniform_ptr(A.W + f * P + rowk)), o, w);
                    }
                }
            } else if (col_real && interior_lane) {
                // ghost rows: y contributions only
                if (A.fold_in_kernel) {
                    // kept for the rows they copy (rows 0/1 below, ny-1/ny-2 after the march)
                    const int layer = (r == -1 || r == ny) ? 0 : 1;
                    gyb[layer][0][lane] = tot.r;
                    gyb[layer][1][lane] = tot.u;
                    gyb[layer][2][lane] = tot.v;
                    gyb[layer][3][lane] = tot.p;
                } else {
                    if (r == -1) ghost_store(A, 2, 0, c, tot);
                    if (r == -2) ghost_store(A, 2, 1, c, tot);
                    if (r == ny) ghost_store(A, 3, 0, c, tot);
                    if (r == ny + 1) ghost_store(A, 3, 1, c, tot);
                }
            }
        }
        // row j so far: own y part + spill from row j-1, then its x gather
        Q4 acc = q4_add(own_j, rsp_prev);
        if (do_x) acc = q4_add(acc, x_j);

        // advance the window
        acc_prev = acc;
        rsp_prev = rsp_j;
        dqm_j = dqm_n;
        qa = qb;
        qb = qc;
        qc = qd;
        Vb[0] = Vc[0]; Vb[1] = Vc[1]; Vb[2] = Vc[2]; Vb[3] = Vc[3];
        Vc[0] = Vn[0]; Vc[1] = Vn[1]; Vc[2] = Vn[2]; Vc[3] = Vn[3];
    }
    __builtin_amdgcn_s_setprio(0);

    if (fold_hi && col_real && interior_lane) {
        // rows ny, ny+1 were finished after the rows they copy: add their
        // cotangents to W of those rows (stored by this wave: wait for the
        // stores, then read W back past the L1)
        __builtin_amdgcn_s_waitcnt(0);
        for (int t = ny - 1; t >= ny - 2 && t >= 0; --t) {
            Q4 d = q4_zero();
            bool any = false;
            for (int layer = 0; layer < 2; ++layer) {
                if (tgt_hi[layer] != t) continue;
                Q4 e = {gyb[layer][0][lane], gyb[layer][1][lane], gyb[layer][2][lane], gyb[layer][3][lane]};
                if (flip_hi[layer]) e.v = -e.v;
                d = q4_add(d, e);
                any = true;
            }
            if (!any) continue;
            const long long rowk = (long long)t * nx;
            const unsigned o = (unsigned)c * 8u;
            double u4[4];
#pragma unroll
            for (int f = 0; f < 4; ++f) u4[f] = ld_buf(buf_rsrc(uniform_ptr(A.U + f * P + rowk)), o);
            const PrimE q = cons_to_prim_e(u4[0], u4[1], u4[2], u4[3], A.g1);
            double du[4];
            cons_vjp(q.r, q.u, q.v, A.g1, d, du);
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const __amdgpu_buffer_rsrc_t rw = buf_rsrc(uniform_ptr(A.W + f * P + rowk));
                const double w0 = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rw, o, 0, 1));
                st_buf(rw, o, w0 + A.c * du[f]);
            }
        }
    }
}

// Fold the ghost-cell cotangents into the cells they copy:
// W(target) += c * (d prim/d cons)^T (sum of M dq_ghost over the ghosts that
// copy the target), M negating the normal velocity on reflective sides.
// One thread per cell of the boundary bands (columns 0, 1, nx-2, nx-1 and
// rows 0, 1, ny-2, ny-1); every ghost maps into these bands, each band cell
// has exactly one thread (no races), and the four sides are summed in a fixed
// order.
__device__ __forceinline__ int band_index(int b, int n) {
    return b < 2 ? b : n - 4 + b;  // 0, 1, n-2, n-1
}

__device__ __forceinline__ bool band_first(int b, int n) {
    // the first of the four band slots that names this index (small n repeats indices)
    const int k = band_index(b, n);
    if (k < 0 || k >= n) return false;
    for (int e = 0; e < b; ++e)
        if (band_index(e, n) == k) return false;
    return true;
}

static __global__ void adj_fold_kernel(const AdjArgs A) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nx = A.nx, ny = A.ny;
    int col, row;
    if (t < 4 * ny) {
        const int b = t / ny;
        if (!band_first(b, nx)) return;
        col = band_index(b, nx);
        row = t - b * ny;
    } else if (t < 4 * ny + 4 * nx) {
        const int u = t - 4 * ny;
        const int b = u / nx;
        if (!band_first(b, ny)) return;
        row = band_index(b, ny);
        col = u - b * nx;
        // cells of the column bands are done by the column-band threads
        for (int e = 0; e < 4; ++e)
            if (band_index(e, nx) == col) return;
    } else {
        return;
    }
    Q4 d = q4_zero();
    bool any = false;
    for (int side = 0; side < 4; ++side) {
        const bool xs = side < 2;
        const int n = xs ? nx : ny;
        const int target = xs ? col : row;
        const int along = xs ? row : col;
        for (int layer = 0; layer < 2; ++layer) {
            const int k = (side == 0 || side == 2) ? -1 - layer : n + layer;
            bool flip;
            const int src = map_index(k, n, A.bc[xs ? 0 : 2], A.bc[xs ? 1 : 3], flip);
            if (src != target) continue;
            const double* g = ghost_at(A, side, layer, along);
            Q4 e = {g[0], g[1], g[2], g[3]};
            if (flip) {
                if (xs) e.u = -e.u;
                else e.v = -e.v;
            }
            d = q4_add(d, e);
            any = true;
        }
    }
    if (!any) return;
    const long long P = A.plane;
    const long long cell = (long long)row * nx + col;
    const double r = A.U[cell], mx = A.U[P + cell], my = A.U[2 * P + cell], E = A.U[3 * P + cell];
    const PrimE q = cons_to_prim_e(r, mx, my, E, A.g1);
    double du[4];
    cons_vjp(q.r, q.u, q.v, A.g1, d, du);
#pragma unroll
    for (int f = 0; f < 4; ++f) A.W[f * P + cell] += A.c * du[f];
}

}  // namespace gd2
