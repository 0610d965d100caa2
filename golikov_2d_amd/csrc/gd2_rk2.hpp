// gd2_rk2.hpp -- one-pass SSP-RK2 step of the Euler MUSCL scheme (SoA
// conserved state) for gfx950.
//
// The two-stage step U1 = U + dt L(U), U' = U/2 + (U1 + dt L(U1))/2 done by
// two stage launches reads and writes the whole intermediate state U1 through
// HBM. Here one wave computes both stages of its strip in one march up the
// rows: a stage-1 producer turns U^n rows into U1 rows two rows ahead of a
// stage-2 consumer that turns U1 rows into U^{n+1}. U1 lives only in
// registers and LDS, so a step moves one read and one write of the state
// (plus halos) instead of five state transfers.
//
// Columns: lane l holds column c = cbase - 2 + l (cbase = 59 tx). The
// producer forms U1 for lanes 0..62 (columns cbase-2 .. cbase+60) from U^n
// columns cbase-4 .. cbase+62; the consumer updates lanes 2..60 (columns
// cbase .. cbase+58), whose five-point stencils stay inside lanes 0..62.
//
// Rows: a strip [j0, j1) needs U1 rows j0-2 .. j1+1. Rows outside the grid on
// a periodic side are formed by the producer like any other row (its loads
// wrap, so they equal the wrapped rows bit for bit); on transmissive and
// reflective sides they are copies (mirrored) of the nearest U1 rows, formed
// after those rows exist, exactly as the stage kernel reads them back
// through the boundary map. Columns outside the grid are treated the same
// way in the consumer's LDS rows.
//
// Every expression is the one the stage kernel uses (same face functions,
// x_balance / full_balance, out_plain, rk_base, out_base, clamps), so the
// result is bitwise that of the two stage launches. With KEEP the producer
// also stores the U1 rows of its strip (bitwise the first stage launch's
// output), for callers that need the stage state, such as an adjoint sweep.
#pragma once

#include "gd2_euler.hpp"

namespace gd2 {

constexpr int kRk2W = 59;     // columns of U^{n+1} per wave
constexpr int kRk2Row = 68;   // LDS row entries: entry e <-> column cbase - 4 + e

struct Rk2Lds {
    double q[2][4][kRk2Row];  // U^n primitive rows (producer x-faces), row r in slot r & 1
    double u[2][4][kWave];    // U^n own conserved values (producer update)
    double w[2][4][kRk2Row];  // U1 primitive rows (consumer x-faces), row r in slot r & 1
    double h[3][4][kWave];    // H = a U^n + b U1 of the three newest U1 rows, row r in slot (r + 3) % 3
};

__device__ __forceinline__ void lds_put68(double (*ring)[4][kRk2Row], int slot, int e, const PrimE& q) {
    ring[slot][0][e] = q.r;
    ring[slot][1][e] = q.u;
    ring[slot][2][e] = q.v;
    ring[slot][3][e] = q.p;
}

__device__ __forceinline__ PrimE lds_get68(const double (*ring)[4][kRk2Row], int slot, int e) {
    PrimE q;
    q.r = ring[slot][0][e];
    q.u = ring[slot][1][e];
    q.v = ring[slot][2][e];
    q.p = ring[slot][3][e];
    return q;
}

__device__ __forceinline__ bool prim_pos(const PrimE& q) { return q.r > kPosMin && q.p > kPosMin; }

// Returns the wave's maximum of the output wave speed (the same on every
// lane); the caller folds it into the clock. TURN: the issue-priority table
// between the two waves of a SIMD (for launches whose waves run one strip
// each and should finish together; persistent waves take the next strip
// instead and leave it off).
template <int ORDER, int FLUX, bool KEEP, bool TURN = true>
__device__ __forceinline__ double euler_rk2_tile(const EulerArgs& A, int tx, int ty, Rk2Lds& lds) {
    const int lane = threadIdx.x;
    const int cbase = tx * kRk2W;
    const int c = cbase - 2 + lane;
    const int j0 = ty * A.rows;
    const int j1 = min(j0 + A.rows, A.ny);
    __syncthreads();  // a wave may run several tiles: the previous tile's LDS reads are done

    // ---- producer (stage 1) geometry, as in the stage kernel
    bool fx, hfx;
    const int ci = map_index(c, A.nx, A.bc[0], A.bc[1], fx);
    const bool has_halo = lane < 3;
    const int hcol = lane < 2 ? c - 2 : cbase + 62;
    const int hci = map_index(hcol, A.nx, A.bc[0], A.bc[1], hfx);
    const int hslot = lane < 2 ? lane : 66;
    const unsigned oc = (unsigned)ci * 8u;
    const unsigned oh = has_halo ? (unsigned)hci * 8u : kOffNone;

    // ---- consumer (stage 2) geometry
    const bool owns = lane >= 2 && lane <= 60 && c < A.nx;
    const unsigned oo = (unsigned)c * 8u;
    // U1 columns outside the grid on a transmissive / reflective side are
    // copies of the grid column the boundary map names (mirrored u when
    // reflective); only the two next to the grid are ever read
    bool gfx = false;
    const int gsrc = map_index(c, A.nx, A.bc[0], A.bc[1], gfx);
    const bool xghost = ((c == -1 || c == -2) && A.bc[0] != GD2_BC_PERIODIC) ||
                        ((c == A.nx || c == A.nx + 1) && A.bc[1] != GD2_BC_PERIODIC);
    const int gsrc_e = gsrc - cbase + 4;  // LDS entry of the source column
    const bool do_xfix = xghost && gsrc_e >= 2 && gsrc_e <= 64;
    const int src_lane = do_xfix ? gsrc_e - 2 : lane;  // the lane holding the source column
    const bool wave_xfix = __any(do_xfix) != 0;
    const bool in_u1 = lane <= 62;  // lanes whose U1 is formed (lane 63 has no right face)

    // ---- rows
    const bool ghost_lo = j0 == 0 && A.bc[2] != GD2_BC_PERIODIC;
    const bool refl_lo = A.bc[2] == GD2_BC_REFLECTIVE;
    const bool refl_hi = A.bc[3] == GD2_BC_REFLECTIVE;
    // the producer forms every needed U1 row inside the grid (and outside it
    // on periodic sides); strips start at row 0 or at row >= 2, so rows below
    // the grid occur only in the first strip, while rows above it occur in
    // every strip that ends within two rows of the top
    const int r_first = ghost_lo ? 0 : j0 - 2;
    const int r_last = A.bc[3] != GD2_BC_PERIODIC ? min(j1 + 1, A.ny - 1) : j1 + 1;

    const double dt = A.clk ? clock_dt(A.clk, A.clk_slot) : (A.dt_dev ? *A.dt_dev : A.dt);
    const double dx = A.dx, dy = A.dy;
    const double dtdV = dt / (dx * dy);
    const double bdt = A.b * dtdV;

    struct Pending {
        Raw own, halo;
        bool fy;
    };
    auto issue_row = [&](int jl) -> Pending {
        Pending pr;
        const int jj = map_index(jl, A.ny, A.bc[2], A.bc[3], pr.fy);
        const RowPtr rp = row_ptr<kSoA>(A, jj);
        pr.own = load_raw(rp, oc);
        pr.halo = load_raw(rp, oh);
        return pr;
    };
    // convert a pending U^n row; to_lds: publish it in LDS slot (primitive
    // row with halo, own conserved values)
    auto finish_row = [&](const Pending& pr, int slot, bool to_lds, bool& pos) -> PrimE {
        PrimE q = finish_cell<kSoA>(A, pr.own, fx, pr.fy).q;
        bool ok = prim_pos(q);
        if (to_lds) {
            lds_put68(lds.q, slot, lane + 2, q);
#pragma unroll
            for (int f = 0; f < 4; ++f) lds.u[slot][f][lane] = pr.own.a[f];
            if (has_halo) {
                PrimE h = finish_cell<kSoA>(A, pr.halo, hfx, pr.fy).q;
                lds_put68(lds.q, slot, hslot, h);
                ok = ok && prim_pos(h);
            }
        }
        pos = __all(ok) != 0;
        return q;
    };

    // ---- producer prologue: U^n rows r_first-2, r_first-1, r_first; the
    // first pass (r = r_first - 1) only forms G1(r_first - 1/2)
    PrimE pv0, pv1;
    Slope ps0;
    FluxE pGm;
    pGm.fr = pGm.fn = pGm.ft = pGm.fe = 0.0;
    unsigned phist;
    Pending pn;
    {
        const Pending a = issue_row(r_first - 2), b = issue_row(r_first - 1), cc = issue_row(r_first);
        pn = issue_row(r_first + 1);
        bool pa, pb, pc;
        const PrimE vm2 = finish_row(a, 0, false, pa);
        pv0 = finish_row(b, 0, false, pb);
        pv1 = finish_row(cc, r_first & 1, true, pc);
        phist = (pa ? 4u : 0u) | (pb ? 2u : 0u) | (pc ? 1u : 0u);
        ps0 = slope3<ORDER>(vm2, pv0, pv1);
    }

    // ---- consumer state: U1 rows j, j+1 (primitive), slope of row j, G2(j-1/2)
    PrimE cw0, cw1;
    cw0.r = cw0.u = cw0.v = cw0.p = 0.0;
    cw1 = cw0;
    Slope cs0;
    cs0.r = cs0.u = cs0.v = cs0.p = 0.0;
    FluxE cGm;
    cGm.fr = cGm.fn = cGm.ft = cGm.fe = 0.0;
    unsigned chist = 0u;  // positivity of the newest U1 rows, newest in bit 0

    double wmax = 0.0;
    double smax;

    // ================= producer pass r: x-face F1 of row r (unless `lead`),
    // the next U^n row into LDS, y-face G1(r+1/2), and U1 row r (unless `lead`)
    // with its H = a U^n + b U1 in LDS
    auto produce = [&](int r, bool lead, PrimE& q1) {
        FluxE Fl, Fr;
        Bal X;
        double smx;
        Raw uc;
        if (!lead) {
            // x-face F1(c-1/2) of row r from LDS; F1(c+1/2) from the next lane
            const int sl = r & 1;
            const PrimE qa = lds_get68(lds.q, sl, lane);
            const PrimE qb = lds_get68(lds.q, sl, lane + 1);
            const PrimE qd = lds_get68(lds.q, sl, lane + 3);
            const PrimE& qc = pv0;
            const Slope sb = slope3<ORDER>(qa, qb, qc);
            const Slope sc = slope3<ORDER>(qb, qc, qd);
            if ((phist & 2u) != 0u) {
                const PrimE L = face_plus<false>(qb, sb), R = face_minus<false>(qc, sc);
                Fl = face_flux<FLUX, false>(A, L.r, L.u, L.v, L.p, R.r, R.u, R.v, R.p, smx);
            } else {
                const PrimE L = face_plus(qb, sb), R = face_minus(qc, sc);
                Fl = face_flux<FLUX>(A, L.r, L.u, L.v, L.p, R.r, R.u, R.v, R.p, smx);
            }
            wall_face_fix<FLUX>(Fl, c == 0 && A.bc[0] == GD2_BC_REFLECTIVE);
            Fr.fr = from_next_lane(Fl.fr);
            Fr.fn = from_next_lane(Fl.fn);
            Fr.ft = from_next_lane(Fl.ft);
            Fr.fe = from_next_lane(Fl.fe);
            X = x_balance(Fl, Fr, dy);
#pragma unroll
            for (int f = 0; f < 4; ++f) uc.a[f] = lds.u[r & 1][f][lane];
        }
        __syncthreads();  // slot r & 1 is read: row r+2 may take it
        bool npos;
        const PrimE pv2 = finish_row(pn, r & 1, true, npos);
        phist = ((phist << 1) | (npos ? 1u : 0u)) & 15u;
        pn = issue_row(r + 3);  // in flight during the rest of the pass
        const Slope ps1 = slope3<ORDER>(pv0, pv1, pv2);
        FluxE Gp;
        if (phist == 15u) {
            const PrimE L = face_plus<false>(pv0, ps0), R = face_minus<false>(pv1, ps1);
            Gp = face_flux<FLUX, false>(A, L.r, L.v, L.u, L.p, R.r, R.v, R.u, R.p, smx);
        } else {
            const PrimE L = face_plus(pv0, ps0), R = face_minus(pv1, ps1);
            Gp = face_flux<FLUX>(A, L.r, L.v, L.u, L.p, R.r, R.v, R.u, R.p, smx);
        }
        wall_face_fix<FLUX>(Gp, r == -1 && A.bc[2] == GD2_BC_REFLECTIVE);
        if (!lead) {
            const Bal P = full_balance(X, pGm, Gp, dx);
            double o0 = out_plain(uc.a[0], dtdV, P.r);
            double o1 = out_plain(uc.a[1], dtdV, P.mx);
            double o2 = out_plain(uc.a[2], dtdV, P.my);
            double o3 = out_plain(uc.a[3], dtdV, P.e);
            // every U1 cell is counted once: by the wave that updates it
            clamp_plain(A, o0, o1, o2, o3, owns && r >= j0 && r < j1);
            if constexpr (KEEP) {
                if (owns && r >= j0 && r < j1) {
                    double* sb = A.Ustage + (long long)r * A.nx;
                    st_buf(buf_rsrc(uniform_ptr(sb)), oo, o0);
                    st_buf(buf_rsrc(uniform_ptr(sb + A.plane)), oo, o1);
                    st_buf(buf_rsrc(uniform_ptr(sb + 2 * A.plane)), oo, o2);
                    st_buf(buf_rsrc(uniform_ptr(sb + 3 * A.plane)), oo, o3);
                }
            }
            q1 = cons_to_prim_e(o0, o1, o2, o3, A.g1);
            const int hs = (r + 3) % 3;
            lds.h[hs][0][lane] = rk_base(A.a, uc.a[0], A.b, o0);
            lds.h[hs][1][lane] = rk_base(A.a, uc.a[1], A.b, o1);
            lds.h[hs][2][lane] = rk_base(A.a, uc.a[2], A.b, o2);
            lds.h[hs][3][lane] = rk_base(A.a, uc.a[3], A.b, o3);
        }
        pv0 = pv1;
        pv1 = pv2;
        ps0 = ps1;
        pGm = Gp;
    };

    // ================= consumer x-face F2 of row j (LDS slot j & 1) and H_j.
    // Only lanes 2..61 need their face F2(c-1/2), so the slope of column c-1
    // is lane l-1's slope of its own column, shifted over (bit for bit).
    auto consume_xface = [&](int j, Bal& X2, double (&hj)[4]) {
        const int sl = j & 1;
        const PrimE qb = lds_get68(lds.w, sl, lane + 1);
        const PrimE qc = lds_get68(lds.w, sl, lane + 2);
        const PrimE qd = lds_get68(lds.w, sl, lane + 3);
        const Slope sc = slope3<ORDER>(qb, qc, qd);
        Slope sb;
        sb.r = from_prev_lane(sc.r);
        sb.u = from_prev_lane(sc.u);
        sb.v = from_prev_lane(sc.v);
        sb.p = from_prev_lane(sc.p);
        FluxE Fl, Fr;
        if ((chist & 2u) != 0u) {
            const PrimE L = face_plus<false>(qb, sb), R = face_minus<false>(qc, sc);
            Fl = face_flux<FLUX, false>(A, L.r, L.u, L.v, L.p, R.r, R.u, R.v, R.p, smax);
        } else {
            const PrimE L = face_plus(qb, sb), R = face_minus(qc, sc);
            Fl = face_flux<FLUX>(A, L.r, L.u, L.v, L.p, R.r, R.u, R.v, R.p, smax);
        }
        wall_face_fix<FLUX>(Fl, c == 0 && A.bc[0] == GD2_BC_REFLECTIVE);
        Fr.fr = from_next_lane(Fl.fr);
        Fr.fn = from_next_lane(Fl.fn);
        Fr.ft = from_next_lane(Fl.ft);
        Fr.fe = from_next_lane(Fl.fe);
        X2 = x_balance(Fl, Fr, dy);
        const int hs = (j + 3) % 3;
#pragma unroll
        for (int f = 0; f < 4; ++f) hj[f] = lds.h[hs][f][lane];
    };

    // ================= consumer: publish U1 row r in LDS slot r & 1 (after
    // the x-face of row r-2 read that slot); columns outside the grid take the
    // (mirrored) value of their source column, in LDS and in the lane's own
    // registers
    auto publish = [&](int r, PrimE& q1) {
        if (wave_xfix) {
            // ghost lanes take their source column's value straight from
            // the lane holding it (a lane permute, no LDS round trip)
            const PrimE sv = {__shfl(q1.r, src_lane), __shfl(q1.u, src_lane), __shfl(q1.v, src_lane),
                              __shfl(q1.p, src_lane)};
            if (do_xfix) {
                q1 = sv;
                if (gfx) q1.u = -q1.u;
            }
        }
        lds_put68(lds.w, r & 1, lane + 2, q1);
    };

    // ================= consumer y-face G2(j+1/2) from rows j-1 .. j+2 (q1 =
    // row j+2), the update of row j when `upd`, and the window advance
    auto consume_y = [&](int j, const PrimE& q1, bool qpos, bool upd, const Bal& X2, const double (&hj)[4]) {
        chist = ((chist << 1) | (qpos ? 1u : 0u)) & 15u;
        const Slope cs1 = slope3<ORDER>(cw0, cw1, q1);
        FluxE Gp;
        if (chist == 15u) {
            const PrimE L = face_plus<false>(cw0, cs0), R = face_minus<false>(cw1, cs1);
            Gp = face_flux<FLUX, false>(A, L.r, L.v, L.u, L.p, R.r, R.v, R.u, R.p, smax);
        } else {
            const PrimE L = face_plus(cw0, cs0), R = face_minus(cw1, cs1);
            Gp = face_flux<FLUX>(A, L.r, L.v, L.u, L.p, R.r, R.v, R.u, R.p, smax);
        }
        wall_face_fix<FLUX>(Gp, j == -1 && A.bc[2] == GD2_BC_REFLECTIVE);
        if (upd && owns) {
            const Bal P = full_balance(X2, cGm, Gp, dx);
            double o0 = out_base(hj[0], bdt, P.r);
            double o1 = out_base(hj[1], bdt, P.mx);
            double o2 = out_base(hj[2], bdt, P.my);
            double o3 = out_base(hj[3], bdt, P.e);
            clamp_wavespeed(A, o0, o1, o2, o3, true, wmax);
            double* ob = A.Uout + (long long)j * A.nx;
            st_buf(buf_rsrc(uniform_ptr(ob)), oo, o0);
            st_buf(buf_rsrc(uniform_ptr(ob + A.plane)), oo, o1);
            st_buf(buf_rsrc(uniform_ptr(ob + 2 * A.plane)), oo, o2);
            st_buf(buf_rsrc(uniform_ptr(ob + 3 * A.plane)), oo, o3);
        }
        cw0 = cw1;
        cw1 = q1;
        cs0 = cs1;
        cGm = Gp;
    };

    // ================= a pass of any kind: U1 rows fed to the consumer are
    // j0-2 .. j1+1; pass r forms U1 row r when the producer owns it and feeds
    // it; the pass before r_first is the producer's lead-in
    auto pass_general = [&](int r) {
        PrimE q1;           // U1 row r, primitive
        bool have = false;  // q1 holds a formed U1 row
        if (r >= r_first - 1 && r <= r_last) {
            const bool lead = r < r_first;
            produce(r, lead, q1);
            have = !lead;
        }
        const int k = r - (j0 - 2);
        if (k < 0) return;
        if (!have) {
            // a row outside the grid on a non-periodic side
            if (r >= A.ny) {
                // top: row ny is row ny-1 (the newest row held), row ny+1 is
                // row ny-1 (transmissive) or row ny-2 (reflective, from LDS)
                if (r == A.ny) {
                    q1 = cw1;
                } else if (!refl_hi) {
                    q1 = cw0;
                } else {
                    bool f;
                    const int src = map_index(r, A.ny, A.bc[2], A.bc[3], f);
                    q1 = lds_get68(lds.w, src & 1, lane + 2);
                }
                if (refl_hi) q1.v = -q1.v;
            } else if (k < 2) {
                return;  // bottom rows -2, -1: formed once rows 0 and 1 exist (k = 2, 3)
            }
        }
        const int j = r - 2;
        const bool upd = k >= 4;
        Bal X2;
        double hj[4];
        if (upd) consume_xface(j, X2, hj);
        __syncthreads();
        if (have) publish(r, q1);
        const bool qpos = __all(!in_u1 || prim_pos(q1)) != 0;
        if (k < 2) {  // rows j0-2, j0-1
            if (k == 0)
                cw0 = q1;
            else
                cw1 = q1;
            chist = ((chist << 1) | (qpos ? 1u : 0u)) & 15u;
            return;
        }
        if (k == 2) {  // row j0
            if (ghost_lo) {
                // row -1 is row 0 (mirrored v when reflective); the slope of
                // row -1 waits for row 1 (reflective row -2 is row 1)
                cw0 = q1;
                if (refl_lo) cw0.v = -cw0.v;
                cw1 = q1;
                chist = qpos ? 3u : 0u;
            } else {
                cs0 = slope3<ORDER>(cw0, cw1, q1);
                cw0 = cw1;
                cw1 = q1;
                chist = ((chist << 1) | (qpos ? 1u : 0u)) & 15u;
            }
            return;
        }
        if (k == 3 && ghost_lo) {
            PrimE m2 = refl_lo ? q1 : cw1;  // row -2: row 1 mirrored, or row 0
            if (refl_lo) m2.v = -m2.v;
            cs0 = slope3<ORDER>(m2, cw0, cw1);
        }
        consume_y(j, q1, qpos, upd, X2, hj);  // k == 3: the lead-in, G2(j0-1/2) only
    };

    // ================= steady pass r: the producer forms U1 row r inside the
    // grid and the consumer updates row r - 2
    auto pass_steady = [&](int r) {
        PrimE q1;
        produce(r, false, q1);
        Bal X2;
        double hj[4];
        consume_xface(r - 2, X2, hj);
        __syncthreads();
        publish(r, q1);
        const bool qpos = __all(!in_u1 || prim_pos(q1)) != 0;
        consume_y(r - 2, q1, qpos, true, X2, hj);
    };

    // the wave behind on its SIMD takes the higher issue priority, decided
    // every kTurnPeriod passes
    SimdTurn turn;
    if constexpr (TURN) turn.init();
    auto set_turn = [&](int pass) {
        if constexpr (TURN) {
            if ((pass & (kTurnPeriod - 1)) == 0) turn.step(pass / kTurnPeriod);
        }
    };

    const int s_begin = j0 + 2;            // first pass that updates a row
    const int s_end = min(r_last, j1 + 1);  // last pass whose U1 row the producer forms
    int r = min(r_first - 1, j0 - 2);
    while (r <= j1 + 1) {
        if (r == s_begin && s_begin <= s_end) {
            for (; r <= s_end; ++r) {
                set_turn(r);
                pass_steady(r);
            }
            continue;
        }
        set_turn(r);
        pass_general(r);
        ++r;
    }
    if constexpr (TURN) turn.finish();
    return wave_max(wmax);
}

template <int ORDER, int FLUX, bool KEEP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu((FLUX == GD2_FLUX_EXACT) ? 1 : 2))) void euler_rk2_kernel(const EulerArgs A) {
    __shared__ Rk2Lds lds;
    if (A.clk && blockIdx.x == 0 && threadIdx.x == 0) clock_commit(A.clk, A.clk_slot, A.dt_log, A.dt_cap);
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const double w = euler_rk2_tile<ORDER, FLUX, KEEP>(A, tile % A.tiles_x, tile / A.tiles_x, lds);
    if (threadIdx.x == 0 && A.wmax_bits) atomic_max_nonneg(A.wmax_bits, w);
}

}  // namespace gd2
