/*
 * gd2_oracle.c -- CPU restatement of the reference's explicit finite-volume
 * step. TEST INFRASTRUCTURE ONLY (the checker of the parity tests, the smoke
 * check and the CPU-baseline leg of bench.py); never part of the product.
 *
 * Built with -O3 -ffp-contract=off (no fused multiply-add, no fast-math) so the
 * arithmetic is the expression-by-expression IEEE evaluation of the reference
 * source lines cited below. The reference's CUDA math (sqrt, pow, fabs, fmin,
 * fmax) maps to the C library.
 *
 * Where the reference has problem-specific pieces (inner sphere, inflow
 * faces, axisymmetric 1/y sources, the hx/hy HLL switch that is never true on
 * its domain) the Cartesian, boundary-kind-driven form is restated instead;
 * each place says so.
 */
#include "gd2_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS  (1e-10)  /* kernel.cu:24 */
#define EPS8 (1e-8)   /* kernel.cu:25 */
#define CPI4 12.56637061435917295384   /* Header.h:20 */
#define CPI8 25.13274122871834590768   /* Header.h:21 */
#define SPI4 (3.544907701811032)       /* Header.h:22 */

/* Iteration caps. The reference loops without a bound (kernel.cu:733-761,
 * 796-826); the caps are far above what any finite input needs and each hit is
 * counted. */
#define NEWTON_CAP 200
#define BISECT_CAP 400

static o_diag g_diag;

void o_diag_reset(void) { memset(&g_diag, 0, sizeof g_diag); }
void o_diag_get(o_diag* out) { *out = g_diag; }

static void diag_inc(long* c)
{
#pragma omp atomic
    (*c)++;
}

int o_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void o_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* pointwise helpers                                                         */
/* ------------------------------------------------------------------------ */

/* kernel.cu:86-100 */
static inline int sgn(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }

/* kernel.cu:51-62 */
double o_minmod(double x, double y)
{
    int s = sgn(x) + sgn(y);
    if (s == 0) return 0.0;
    return (s / 2.0) * fmin(fabs(x), fabs(y));
}

/* kernel.cu:69-73 */
double o_linear(double x1, double t1, double x2, double t2, double x3, double t3, double y)
{
    double d = o_minmod((t1 - t2) / (x1 - x2), (t2 - t3) / (x2 - x3));
    return d * (y - x2) + t2;
}

/* kernel.cu:75-84 */
void o_linear2(double x1, double t1, double x2, double t2, double x3, double t3,
               double y1, double y2, double* A, double* B)
{
    double d = o_minmod((t1 - t2) / (x1 - x2), (t2 - t3) / (x2 - x3));
    *A = d * (y1 - x2) + t2;
    *B = d * (y2 - x2) + t2;
}

/* ------------------------------------------------------------------------ */
/* HLL (Euler), kernel.cu:1485-1609: sound speeds 0 where rho <= 0, bounds   */
/* D_L = min(Vn_L, Vn_R) - max(c_L, c_R), D_R = max(Vn_L, Vn_R) +            */
/* max(c_L, c_R); the left flux when D_L > Omega (= 0), the right one when   */
/* D_R < Omega, else the HLL average; fluxes in the lab frame along n.       */
/* Returns the face time step krit rad / max(|D_L|, |D_R|).                  */
/* ------------------------------------------------------------------------ */
double o_hll(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
             double n1, double n2, o_d2* Ps, o_d2* Pu, double rad)
{
    const double ro_L = Ls.x, p_L = Ls.y, v1_L = Lu.x, v2_L = Lu.y;
    const double ro_R = Rs.x, p_R = Rs.y, v1_R = Ru.x, v2_R = Ru.y;
    const double Vkv_L = v1_L * v1_L + v2_L * v2_L;
    const double Vkv_R = v1_R * v1_R + v2_R * v2_R;
    const double c_L = ro_L <= 0 ? 0.0 : sqrt(gamma * p_L / ro_L);
    const double c_R = ro_R <= 0 ? 0.0 : sqrt(gamma * p_R / ro_R);
    const double e_L = p_L / (gamma - 1.0) + ro_L * Vkv_L / 2.0;
    const double e_R = p_R / (gamma - 1.0) + ro_R * Vkv_R / 2.0;
    const double Vn_L = v1_L * n1 + v2_L * n2;
    const double Vn_R = v1_R * n1 + v2_R * n2;
    const double D_L = fmin(Vn_L, Vn_R) - fmax(c_L, c_R);
    const double D_R = fmax(Vn_L, Vn_R) + fmax(c_L, c_R);
    double t = 10000000;
    t = fmin(t, krit * rad / fmax(fabs(D_L), fabs(D_R)));
    /* lab-frame physical fluxes projected on n */
    const double fl_1 = ro_L * v1_L * n1 + ro_L * v2_L * n2;
    const double fl_2 = (ro_L * v1_L * v1_L + p_L) * n1 + ro_L * v1_L * v2_L * n2;
    const double fl_3 = ro_L * v1_L * v2_L * n1 + (ro_L * v2_L * v2_L + p_L) * n2;
    const double fl_5 = (e_L + p_L) * v1_L * n1 + (e_L + p_L) * v2_L * n2;
    if (D_L > 0.0) {
        Ps->x = fl_1; Pu->x = fl_2; Pu->y = fl_3; Ps->y = fl_5;
        return t;
    }
    const double fr_1 = ro_R * v1_R * n1 + ro_R * v2_R * n2;
    const double fr_2 = (ro_R * v1_R * v1_R + p_R) * n1 + ro_R * v1_R * v2_R * n2;
    const double fr_3 = ro_R * v1_R * v2_R * n1 + (ro_R * v2_R * v2_R + p_R) * n2;
    const double fr_5 = (e_R + p_R) * v1_R * n1 + (e_R + p_R) * v2_R * n2;
    if (D_R < 0.0) {
        Ps->x = fr_1; Pu->x = fr_2; Pu->y = fr_3; Ps->y = fr_5;
        return t;
    }
    const double dq1 = ro_R - ro_L;
    const double dq2 = ro_R * v1_R - ro_L * v1_L;
    const double dq3 = ro_R * v2_R - ro_L * v2_L;
    const double dq5 = e_R - e_L;
    Ps->x = (D_R * fl_1 - D_L * fr_1 + D_L * D_R * dq1) / (D_R - D_L);
    Pu->x = (D_R * fl_2 - D_L * fr_2 + D_L * D_R * dq2) / (D_R - D_L);
    Pu->y = (D_R * fl_3 - D_L * fr_3 + D_L * D_R * dq3) / (D_R - D_L);
    Ps->y = (D_R * fl_5 - D_L * fr_5 + D_L * D_R * dq5) / (D_R - D_L);
    return t;
}

/* ------------------------------------------------------------------------ */
/* HLLC (Euler): HLLC_Aleksashov, kernel.cu:853-1165, called with B = 0.     */
/* With B == 0 the frame of kernel.cu:915-945 is (n, e_z, t) with            */
/* t = (-n2, n1), every magnetic term is an exact zero, the fast speed       */
/* (qp+qm)/2 equals c exactly, and the |UZ[5]| < eps branch (1095-1105)      */
/* keeps the tangential velocities; the lines below keep the reference's     */
/* operation order for every term that survives.                             */
/* ------------------------------------------------------------------------ */
double o_hllc(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
              double n1, double n2, o_d2* Ps, o_d2* Pu, double rad)
{
    const double g1 = gamma - 1.0;
    const double t1 = -n2, t2 = n1;           /* aco[.][2], kernel.cu:942-944 */
    double r1 = Ls.x, u1 = Lu.x, v1 = Lu.y, p1 = Ls.y;
    double r2 = Rs.x, u2 = Ru.x, v2 = Ru.y, p2 = Rs.y;

    /* kernel.cu:947-953: normal and tangential components */
    double vL0 = n1 * u1 + n2 * v1, vL2 = t1 * u1 + t2 * v1;
    double vR0 = n1 * u2 + n2 * v2, vR2 = t1 * u2 + t2 * v2;

    /* kernel.cu:955-971 with B = 0 */
    double cL = sqrt(gamma * p1 / r1);
    double cR = sqrt(gamma * p2 / r2);
    double ptL = p1, ptR = p2;

    /* kernel.cu:981-986 */
    double SL = fmin(vL0 - cL, vR0 - cR);
    double SR = fmax(vL0 + cL, vR0 + cR);
    double suR = SR - vR0;
    double suL = SL - vL0;
    double SM = (suR * r2 * vR0 - ptR + ptL - suL * r1 * vL0) / (suR * r2 - suL * r1);

    /* kernel.cu:1014-1015 */
    double UU = fmax(fabs(SL), fabs(SR));
    double time = krit * rad / UU;

    /* kernel.cu:1017-1024 */
    double upt1 = (u1 * u1 + v1 * v1) / 2.0;
    double upt2 = (u2 * u2 + v2 * v2) / 2.0;
    double e1 = p1 / g1 + r1 * upt1;
    double e2 = p2 / g1 + r2 * upt2;

    /* kernel.cu:1026-1047: FL[0,1,3,4], UL[0,1,3,4] (FL[2] = UL[2] = 0) */
    double FL0 = r1 * vL0, FL1 = r1 * vL0 * vL0 + ptL, FL3 = r1 * vL0 * vL2, FL4 = (e1 + ptL) * vL0;
    double FR0 = r2 * vR0, FR1 = r2 * vR0 * vR0 + ptR, FR3 = r2 * vR0 * vR2, FR4 = (e2 + ptR) * vR0;
    double UL1 = r1 * vL0, UL3 = r1 * vL2;
    double UR1 = r2 * vR0, UR3 = r2 * vR2;

    /* kernel.cu:1063-1093 */
    double suRm = suR / (SR - SM);
    double suLm = suL / (SL - SM);
    double rzR = r2 * suRm;
    double rzL = r1 * suLm;
    double ptzR = ptR + r2 * suR * (SM - vR0);
    double ptzL = ptL + r1 * suL * (SM - vL0);
    double ptz = (ptzR + ptzL) / 2.0;
    double ezR = e2 * suRm + (ptz * SM - ptR * vR0) / (SR - SM);
    double ezL = e1 * suLm + (ptz * SM - ptL * vL0) / (SL - SM);

    double q0, q2;
    /* kernel.cu:1119-1158 (wv = 0) */
    if (SL > 0.0) {
        Ps->x = FL0; Ps->y = FL4; q0 = FL1; q2 = FL3;
    } else if (SL <= 0.0 && SM >= 0.0) {
        Ps->x = FL0 + SL * (rzL - r1);
        Ps->y = FL4 + SL * (ezL - e1);
        q0 = FL1 + SL * (SM * rzL - UL1);
        q2 = FL3 + SL * (vL2 * rzL - UL3);
    } else if (SM <= 0.0 && SR >= 0.0) {
        Ps->x = FR0 + SR * (rzR - r2);
        Ps->y = FR4 + SR * (ezR - e2);
        q0 = FR1 + SR * (SM * rzR - UR1);
        q2 = FR3 + SR * (vR2 * rzR - UR3);
    } else if (SR < 0.0) {
        Ps->x = FR0; Ps->y = FR4; q0 = FR1; q2 = FR3;
    } else {
        /* kernel.cu:1155-1158: no branch (NaN input); the reference leaves the
         * flux uninitialised, the restatement returns NaN. */
        diag_inc(&g_diag.bad_branch);
        Ps->x = Ps->y = NAN; q0 = q2 = NAN;
    }
    /* kernel.cu:1161-1162 */
    Pu->x = n1 * q0 + t1 * q2;
    Pu->y = n2 * q0 + t2 * q2;
    return time;
}

/* ------------------------------------------------------------------------ */
/* HLLC_Aleksashov2, kernel.cu:1167-1348. The HLLC of HLLC_Aleksashov with   */
/* B = 0 in the frame (n, t = (-n2, n1)), except for its contact speed: the  */
/* numerator takes the left state's TANGENTIAL velocity vL[1] in the term    */
/* suL r1 vL[1] (kernel.cu:1222), and SM = 0 when the denominator is 0       */
/* (kernel.cu:1219-1223). Restated as written.                               */
/* ------------------------------------------------------------------------ */
double o_hllc_aleksashov2(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
                          double n1, double n2, o_d2* Ps, o_d2* Pu, double rad)
{
    const double g1 = gamma - 1.0;
    double r1 = Ls.x, p1 = Ls.y, u1 = Lu.x, v1 = Lu.y;
    double r2 = Rs.x, p2 = Rs.y, u2 = Ru.x, v2 = Ru.y;
    /* kernel.cu:1187-1201: aco = [[n1, -n2], [n2, n1]] */
    double vL0 = n1 * u1 + n2 * v1, vL1 = -n2 * u1 + n1 * v1;
    double vR0 = n1 * u2 + n2 * v2, vR1 = -n2 * u2 + n1 * v2;
    /* kernel.cu:1208-1217 */
    double cL = sqrt(gamma * p1 / r1);
    double cR = sqrt(gamma * p2 / r2);
    double SL = fmin(vL0 - cL, vR0 - cR);
    double SR = fmax(vL0 + cL, vR0 + cR);
    double t = 10000000;
    t = fmin(t, krit * rad / fmax(fabs(SL), fabs(SR)));
    /* kernel.cu:1219-1225 */
    double suR = SR - vR0;
    double suL = SL - vL0;
    double SM = 0.0;
    if (fabs(suR * r2 - suL * r1) > 0) SM = (suR * r2 * vR0 - p2 + p1 - suL * r1 * vL1) / (suR * r2 - suL * r1);
    /* kernel.cu:1232-1261 */
    double upt1 = (u1 * u1 + v1 * v1) / 2.0;
    double upt2 = (u2 * u2 + v2 * v2) / 2.0;
    double e1 = p1 / g1 + r1 * upt1;
    double e2 = p2 / g1 + r2 * upt2;
    double FL0 = r1 * vL0, FL1 = r1 * vL0 * vL0 + p1, FL2 = r1 * vL0 * vL1, FL3 = (e1 + p1) * vL0;
    double FR0 = r2 * vR0, FR1 = r2 * vR0 * vR0 + p2, FR2 = r2 * vR0 * vR1, FR3 = (e2 + p2) * vR0;
    double UL1 = r1 * vL0, UL2 = r1 * vL1, UR1 = r2 * vR0, UR2 = r2 * vR1;
    /* kernel.cu:1264-1297 */
    double suRm = suR / (SR - SM);
    double suLm = suL / (SL - SM);
    double rzR = r2 * suRm;
    double rzL = r1 * suLm;
    double ptzR = p2 + r2 * suR * (SM - vR0);
    double ptzL = p1 + r1 * suL * (SM - vL0);
    double ptz = (ptzR + ptzL) / 2.0;
    double ezR = e2 * suRm + (ptz * SM - p2 * vR0) / (SR - SM);
    double ezL = e1 * suLm + (ptz * SM - p1 * vL0) / (SL - SM);
    double q0, q1;
    /* kernel.cu:1301-1341 */
    if (SL > 0.0) {
        Ps->x = FL0; Ps->y = FL3; q0 = FL1; q1 = FL2;
    } else if (SL <= 0.0 && SM >= 0.0) {
        Ps->x = FL0 + SL * (rzL - r1);
        Ps->y = FL3 + SL * (ezL - e1);
        q0 = FL1 + SL * (SM * rzL - UL1);
        q1 = FL2 + SL * (vL1 * rzL - UL2);
    } else if (SM <= 0.0 && SR >= 0.0) {
        Ps->x = FR0 + SR * (rzR - r2);
        Ps->y = FR3 + SR * (ezR - e2);
        q0 = FR1 + SR * (SM * rzR - UR1);
        q1 = FR2 + SR * (vR1 * rzR - UR2);
    } else if (SR < 0.0) {
        Ps->x = FR0; Ps->y = FR3; q0 = FR1; q1 = FR2;
    } else {
        diag_inc(&g_diag.bad_branch);
        Ps->x = Ps->y = NAN; q0 = q1 = NAN;
    }
    /* kernel.cu:1343-1344 */
    Pu->x = n1 * q0 + -n2 * q1;
    Pu->y = n2 * q0 + n1 * q1;
    return t;
}

/* ------------------------------------------------------------------------ */
/* HLLC_Korolkov, kernel.cu:1349-1483: lab-frame fluxes projected on n; the  */
/* outer speeds of the HLL solver (D_L = min(Vn) - max(c), D_R = max(Vn) +   */
/* max(c)); contact speed D_C and star pressure P_T of the two-state HLLC;   */
/* the star state takes the normal velocity D_C, keeps the tangential one,   */
/* and its energy from the equation of state, P_T/(gamma-1) + rho*|V*|^2/2   */
/* (kernel.cu:1448, 1468). Omega = 0 (kernel.cu:11).                         */
/* ------------------------------------------------------------------------ */
double o_hllc_korolkov(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
                       double n1, double n2, o_d2* Ps, o_d2* Pu, double rad)
{
    const double ro_L = Ls.x, p_L = Ls.y, v1_L = Lu.x, v2_L = Lu.y;
    const double ro_R = Rs.x, p_R = Rs.y, v1_R = Ru.x, v2_R = Ru.y;
    /* kernel.cu:1366-1382 */
    const double Vkv_L = v1_L * v1_L + v2_L * v2_L;
    const double Vkv_R = v1_R * v1_R + v2_R * v2_R;
    const double c_L = sqrt(gamma * p_L / ro_L);
    const double c_R = sqrt(gamma * p_R / ro_R);
    const double e_L = p_L / (gamma - 1.0) + ro_L * Vkv_L / 2.0;
    const double e_R = p_R / (gamma - 1.0) + ro_R * Vkv_R / 2.0;
    const double Vn_L = v1_L * n1 + v2_L * n2;
    const double Vn_R = v1_R * n1 + v2_R * n2;
    const double D_L = fmin(Vn_L, Vn_R) - fmax(c_L, c_R);
    const double D_R = fmax(Vn_L, Vn_R) + fmax(c_L, c_R);
    double t = 10000000;
    t = fmin(t, krit * rad / fmax(fabs(D_L), fabs(D_R)));
    /* kernel.cu:1384-1397 */
    const double fl_1 = ro_L * v1_L * n1 + ro_L * v2_L * n2;
    const double fl_2 = (ro_L * v1_L * v1_L + p_L) * n1 + ro_L * v1_L * v2_L * n2;
    const double fl_3 = ro_L * v1_L * v2_L * n1 + (ro_L * v2_L * v2_L + p_L) * n2;
    const double fl_5 = (e_L + p_L) * v1_L * n1 + (e_L + p_L) * v2_L * n2;
    if (D_L > 0.0) {
        Ps->x = fl_1; Pu->x = fl_2; Pu->y = fl_3; Ps->y = fl_5;
        return t;
    }
    /* kernel.cu:1408-1421 */
    const double fr_1 = ro_R * v1_R * n1 + ro_R * v2_R * n2;
    const double fr_2 = (ro_R * v1_R * v1_R + p_R) * n1 + ro_R * v1_R * v2_R * n2;
    const double fr_3 = ro_R * v1_R * v2_R * n1 + (ro_R * v2_R * v2_R + p_R) * n2;
    const double fr_5 = (e_R + p_R) * v1_R * n1 + (e_R + p_R) * v2_R * n2;
    if (D_R < 0.0) {
        Ps->x = fr_1; Pu->x = fr_2; Pu->y = fr_3; Ps->y = fr_5;
        return t;
    }
    /* kernel.cu:1432-1441 */
    const double u_L = Vn_L, u_R = Vn_R;
    const double D_C = ((D_R - u_R) * ro_R * u_R - (D_L - u_L) * ro_L * u_L - p_R + p_L) /
                       ((D_R - u_R) * ro_R - (D_L - u_L) * ro_L);
    const double roro_L = ro_L * ((D_L - u_L) / (D_L - D_C));
    const double roro_R = ro_R * ((D_R - u_R) / (D_R - D_C));
    const double P_T = (p_L * ro_R * (u_R - D_R) - p_R * ro_L * (u_L - D_L) -
                        ro_L * ro_R * (u_L - D_L) * (u_R - D_R) * (u_R - u_L)) /
                       (ro_R * (u_R - D_R) - ro_L * (u_L - D_L));
    if (D_L <= 0.0 && D_C >= 0.0) {
        /* kernel.cu:1443-1461 */
        const double Vx = v1_L + (D_C - Vn_L) * n1;
        const double Vy = v2_L + (D_C - Vn_L) * n2;
        const double ee_L = P_T / (gamma - 1.0) + roro_L * (Vx * Vx + Vy * Vy) / 2.0;
        Ps->x = D_L * (roro_L - ro_L) + fl_1;
        Pu->x = D_L * (roro_L * Vx - ro_L * v1_L) + fl_2;
        Pu->y = D_L * (roro_L * Vy - ro_L * v2_L) + fl_3;
        Ps->y = D_L * (ee_L - e_L) + fl_5;
    } else if (D_R >= 0.0 && D_C <= 0.0) {
        /* kernel.cu:1463-1480 */
        const double Vx = v1_R + (D_C - Vn_R) * n1;
        const double Vy = v2_R + (D_C - Vn_R) * n2;
        const double ee_R = P_T / (gamma - 1.0) + roro_R * (Vx * Vx + Vy * Vy) / 2.0;
        Ps->x = D_R * (roro_R - ro_R) + fr_1;
        Pu->x = D_R * (roro_R * Vx - ro_R * v1_R) + fr_2;
        Pu->y = D_R * (roro_R * Vy - ro_R * v2_R) + fr_3;
        Ps->y = D_R * (ee_R - e_R) + fr_5;
    } else {
        /* no branch (NaN input): the reference leaves the flux unset */
        diag_inc(&g_diag.bad_branch);
        Ps->x = Ps->y = Pu->x = Pu->y = NAN;
    }
    return t;
}

/* ------------------------------------------------------------------------ */
/* HLLCQ_Korolkov_2D, Gas_Dinamic.cu:359-464 (HLLC_Korolkov_2D, 265-358, is  */
/* the same arithmetic without the scalar: PQ = NULL). Face frame (n, t =     */
/* (-n2, n1)), outer speeds min(uL - cL, uR - cR) / max(uL + cL, uR + cR),    */
/* the one-sided star pressure pp = p1 + rho1 (SL - uL)(SM - uL), the star    */
/* energy from the Rankine-Hugoniot condition, SM <= 0 on the right.          */
/* PQ: flux of the passive scalar Q carried with the density.                 */
/* ------------------------------------------------------------------------ */
double o_hllcq_korolkov_2d(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru, double Q_L,
                           double Q_R, double n1, double n2, o_d2* Ps, o_d2* Pu, double* PQ, double rad)
{
    const double g1 = gamma - 1.0;
    const double ro1 = Ls.x, u1 = Lu.x, v1 = Lu.y, p1 = Ls.y;
    const double ro2 = Rs.x, u2 = Ru.x, v2 = Ru.y, p2 = Rs.y;
    const double t1 = -n2, t2 = n1;
    /* Gas_Dinamic.cu:383-400 */
    const double u_L = u1 * n1 + v1 * n2, v_L = u1 * t1 + v1 * t2;
    const double u_R = u2 * n1 + v2 * n2, v_R = u2 * t1 + v2 * t2;
    const double cL = sqrt(gamma * p1 / ro1);
    const double cR = sqrt(gamma * p2 / ro2);
    const double SL = fmin(u_L - cL, u_R - cR);
    const double SR = fmax(u_L + cL, u_R + cR);
    const double UU = fmax(fabs(SL), fabs(SR));
    const double time = krit * rad / UU;
    /* Gas_Dinamic.cu:405-409 */
    const double FL1 = ro1 * u_L * u_L + p1, FL2 = ro1 * u_L * v_L;
    const double FR1 = ro2 * u_R * u_R + p2, FR2 = ro2 * u_R * v_R;
    double q = 0.0;
    if (SL >= 0.0) {
        /* Gas_Dinamic.cu:411-419 */
        q = Q_L * u_L;
        Ps->x = ro1 * u_L;
        Ps->y = (gamma * p1 / g1 + 0.5 * ro1 * (u1 * u1 + v1 * v1)) * u_L;
        Pu->x = FL1 * n1 + FL2 * t1;
        Pu->y = FL1 * n2 + FL2 * t2;
    } else if (SR <= 0.0) {
        /* Gas_Dinamic.cu:420-428 */
        q = Q_R * u_R;
        Ps->x = ro2 * u_R;
        Ps->y = (gamma * p2 / g1 + 0.5 * ro2 * (u2 * u2 + v2 * v2)) * u_R;
        Pu->x = FR1 * n1 + FR2 * t1;
        Pu->y = FR1 * n2 + FR2 * t2;
    } else {
        /* Gas_Dinamic.cu:429-431 */
        const double SM = ((SR - u_R) * ro2 * u_R - (SL - u_L) * ro1 * u_L - p2 + p1) /
                          ((SR - u_R) * ro2 - (SL - u_L) * ro1);
        const double pp = p1 + ro1 * (SL - u_L) * (SM - u_L);
        if (SM <= 0.0) {
            /* Gas_Dinamic.cu:433-447 */
            const double rr = ro2 * (SR - u_R) / (SR - SM);
            const double e = p2 / g1 + 0.5 * ro2 * (u2 * u2 + v2 * v2);
            const double ee = ((SR - u_R) * e - p2 * u_R + pp * SM) / (SR - SM);
            q = SR * (rr * Q_R / ro2 - Q_R) + Q_R * u_R;
            Ps->x = SR * (rr - ro2) + ro2 * u_R;
            Ps->y = SR * (ee - e) + (gamma * p2 / g1 + 0.5 * ro2 * (u2 * u2 + v2 * v2)) * u_R;
            const double F1 = FR1 + SR * (rr * SM - ro2 * u_R);
            const double F2 = FR2 + SR * (rr * v_R - ro2 * v_R);
            Pu->x = F1 * n1 + F2 * t1;
            Pu->y = F1 * n2 + F2 * t2;
        } else if (SM >= 0.0) {
            /* Gas_Dinamic.cu:449-462 */
            const double rr = ro1 * (SL - u_L) / (SL - SM);
            const double e = p1 / g1 + 0.5 * ro1 * (u1 * u1 + v1 * v1);
            const double ee = ((SL - u_L) * e - p1 * u_L + pp * SM) / (SL - SM);
            q = SL * (rr * Q_L / ro1 - Q_L) + Q_L * u_L;
            Ps->x = SL * (rr - ro1) + ro1 * u_L;
            Ps->y = SL * (ee - e) + (gamma * p1 / g1 + 0.5 * ro1 * (u1 * u1 + v1 * v1)) * u_L;
            const double F1 = FL1 + SL * (rr * SM - ro1 * u_L);
            const double F2 = FL2 + SL * (rr * v_L - ro1 * v_L);
            Pu->x = F1 * n1 + F2 * t1;
            Pu->y = F1 * n2 + F2 * t2;
        } else {
            diag_inc(&g_diag.bad_branch);
            q = NAN;
            Ps->x = Ps->y = Pu->x = Pu->y = NAN;
        }
    }
    if (PQ) *PQ = q;
    return time;
}

/* ------------------------------------------------------------------------ */
/* HLLCQ_Aleksashov, Gas_Dinamic.cu:465-785, as the reference calls it: B = 0 */
/* (bx1..bz2 = 0, Gas_Dinamic.cu:484-499). Then d = 0 picks the frame        */
/* (n, e_z, (-n2, n1)) (Gas_Dinamic.cu:525-552), the fast speeds equal c     */
/* (sqrt(c*c) = c in IEEE arithmetic), every magnetic term is an exact zero, */
/* the repeated speed block (Gas_Dinamic.cu:592-608) recomputes the same     */
/* values, and |UZ[5]| < 1e-9 keeps the tangential velocities: the gas       */
/* fluxes are HLLC_Aleksashov's (o_hllc) term for term. PQ (Gas_Dinamic.cu:   */
/* 740-772) carries the passive scalar with the density; wv = 0.             */
/* ------------------------------------------------------------------------ */
double o_hllcq_aleksashov(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru, double Q_L,
                          double Q_R, double n1, double n2, o_d2* Ps, o_d2* Pu, double* PQ, double rad)
{
    const double time = o_hllc(gamma, krit, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
    /* the scalar's branch quantities, as in Gas_Dinamic.cu:569-579, 676-686 */
    const double r1 = Ls.x, p1 = Ls.y, r2 = Rs.x, p2 = Rs.y;
    const double vL0 = n1 * Lu.x + n2 * Lu.y, vR0 = n1 * Ru.x + n2 * Ru.y;
    const double cL = sqrt(gamma * p1 / r1), cR = sqrt(gamma * p2 / r2);
    const double SL = fmin(vL0 - cL, vR0 - cR);
    const double SR = fmax(vL0 + cL, vR0 + cR);
    const double suR = SR - vR0, suL = SL - vL0;
    const double SM = (suR * r2 * vR0 - p2 + p1 - suL * r1 * vL0) / (suR * r2 - suL * r1);
    const double rzR = r2 * (suR / (SR - SM)), rzL = r1 * (suL / (SL - SM));
    const double FL0 = Q_L * vL0, FR0 = Q_R * vR0;
    double q;
    if (SL > 0.0) q = FL0;
    else if (SL <= 0.0 && SM >= 0.0) q = FL0 + SL * (rzL * Q_L / r1 - Q_L);
    else if (SM <= 0.0 && SR >= 0.0) q = FR0 + SR * (rzR * Q_R / r2 - Q_R);
    else if (SR < 0.0) q = FR0;
    else q = NAN;
    if (PQ) *PQ = q;
    return time;
}

/* ------------------------------------------------------------------------ */
/* Exact Riemann solver: Godunov_Solver_Alexashov + newton + devtwo + lev     */
/* (kernel.cu:214-656, 721-851), gamma as a runtime value.                    */
/* ------------------------------------------------------------------------ */
typedef struct { double ga, g1, g2, gp, gm; } gconst;

static gconst make_gconst(double gamma)
{
    gconst c;
    c.ga = gamma;
    c.g1 = gamma - 1.0;                /* Header.h:6 */
    c.g2 = gamma + 1.0;                /* Header.h:8 */
    c.gp = (c.g2 / c.ga) / 2.0;        /* Header.h:10 */
    c.gm = (c.g1 / c.ga) / 2.0;        /* Header.h:11 */
    return c;
}

/* kernel.cu:837-851 */
static void lev(const gconst* k, double enI, double pI, double rI, double enII, double pII,
                double rII, double uuu, double* fee)
{
    double cI = sqrt(k->ga * pI / rI);
    double cII = sqrt(k->ga * pII / rII);
    double fI = (uuu - pI) / (rI * cI * sqrt(k->gp * uuu / pI + k->gm));
    double fII = 2.0 / k->g1 * cII * (pow((uuu / pII), k->gm) - 1.0);
    double f1 = fI + fII;
    double f2 = enI - enII;
    *fee = f1 - f2;
}

/* kernel.cu:768-835 (bisection on [pI, pII], stop when |f| <= 1e-9) */
static double devtwo(const gconst* k, double enI, double pI, double rI, double enII,
                     double pII, double rII)
{
    const double epsil = 10e-10;
    double kl = pI, kp = pII, kc, ksi, ksir, ksit;
    lev(k, enI, pI, rI, enII, pII, rII, kl, &ksi);
    lev(k, enI, pI, rI, enII, pII, rII, kp, &ksir);
    if (fabs(ksi) <= epsil) return kl;
    if (fabs(ksir) <= epsil) return kp;
    for (int it = 1;; ++it) {
        kc = (kl + kp) / 2.0;
        lev(k, enI, pI, rI, enII, pII, rII, kc, &ksit);
        if (fabs(ksit) <= epsil) return kc;
        if (it >= BISECT_CAP) { diag_inc(&g_diag.bisect_cap_hits); return kc; }
        if ((ksi * ksit) <= 0.0) { kp = kc; ksir = ksit; }
        else { kl = kc; ksi = ksit; }
    }
}

/* kernel.cu:721-766 (Newton on the two-shock pressure function) */
static double newton(const gconst* k, double enI, double pI, double rI, double enII,
                     double pII, double rII)
{
    double cI = sqrt(k->ga * pI / rI);
    double cII = sqrt(k->ga * pII / rII);
    double pn = pI * rII * cII + pII * rI * cI + (enI - enII) * rI * cI * rII * cII;
    pn = pn / (rI * cI + rII * cII);
    double pee = pn;
    double p;
    for (int kiter = 1;; ++kiter) {
        p = pn;
        double fI = (p - pI) / (rI * cI * sqrt(k->gp * p / pI + k->gm));
        double fIs = (k->ga + 1.0) * p / pI + (3.0 * k->ga - 1.0);
        fIs = fIs / (4.0 * k->ga * rI * cI * pow((k->gp * p / pI + k->gm), (3.0 / 2.0)));
        double fII = (p - pII) / (rII * cII * sqrt(k->gp * p / pII + k->gm));
        double fIIs = (k->ga + 1.0) * p / pII + (3.0 * k->ga - 1.0);
        fIIs = fIIs / (4.0 * k->ga * rII * cII * pow((k->gp * p / pII + k->gm), (3.0 / 2.0)));
        pn = p - (fI + fII - (enI - enII)) / (fIs + fIIs);
        if (!(fabs(pn / pee - p / pee) >= EPS)) break;
        if (kiter >= NEWTON_CAP) { diag_inc(&g_diag.newton_cap_hits); break; }
    }
    return pn;
}

/* Wave-pattern selection and star pressure (kernel.cu:230-317). Outputs the
 * swapped-frame variables; il and ip as in the reference (-1 vacuum). */
typedef struct {
    double enI, teI2, pI, rI, enII, teII2, pII, rII, cI, cII, p;
    int il, ip, ipiz;
} gstate;

static void godunov_setup(const gconst* k, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
                          double n1, double n2, gstate* g)
{
    double al = n1, be = n2, al2 = -n2, be2 = n1;
    g->enI = al * Lu.x + be * Lu.y;
    g->teI2 = al2 * Lu.x + be2 * Lu.y;
    g->enII = al * Ru.x + be * Ru.y;
    g->teII2 = al2 * Ru.x + be2 * Ru.y;
    g->pI = Ls.y; g->pII = Rs.y; g->rI = Ls.x; g->rII = Rs.x;
    g->ipiz = 0;
    if (g->pI > g->pII) {   /* kernel.cu:243-270 */
        double eno2 = g->enII, teo22 = g->teII2, p2 = g->pII, r2 = g->rII;
        double eno1 = g->enI, teo12 = g->teI2, p1 = g->pI, r1 = g->rI;
        g->enI = -eno2; g->teI2 = teo22; g->pI = p2; g->rI = r2;
        g->enII = -eno1; g->teII2 = teo12; g->pII = p1; g->rII = r1;
        g->ipiz = 1;
    }
    g->cI = 0.0; g->cII = 0.0;   /* kernel.cu:272-281 */
    if (g->rI != 0.0) g->cI = sqrt(k->ga * g->pI / g->rI);
    if (g->rII != 0.0) g->cII = sqrt(k->ga * g->pII / g->rII);

    /* kernel.cu:286-317 */
    double a = sqrt(g->rI * (k->g2 * g->pII + k->g1 * g->pI) / 2.0);
    double Uud = (g->pII - g->pI) / a;
    double Urz = -2.0 * g->cII / k->g1 * (1.0 - pow((g->pI / g->pII), k->gm));
    double Uvk = -2.0 * (g->cII + g->cI) / k->g1;
    double Udf = g->enI - g->enII;
    g->p = 0.0;
    g->il = g->ip = -2;   /* the reference leaves them unset for NaN input */
    if (Udf < Uvk) {
        g->il = -1; g->ip = -1;
    } else if ((Udf >= Uvk) && (Udf <= Urz)) {
        g->p = g->pI * pow(((Udf - Uvk) / (Urz - Uvk)), (1.0 / k->gm));
        g->il = 0; g->ip = 0;
    } else if ((Udf > Urz) && (Udf <= Uud)) {
        g->p = devtwo(k, g->enI, g->pI, g->rI, g->enII, g->pII, g->rII);
        g->il = 1; g->ip = 0;
    } else if (Udf > Uud) {
        g->p = newton(k, g->enI, g->pI, g->rI, g->enII, g->pII, g->rII);
        g->il = 1; g->ip = 1;
    }
}

/* expansion-fan coefficient a = gm*r*c*(1-p/pK)/(1-(p/pK)^gm), kernel.cu:383-390 */
static inline double fan_a(const gconst* k, double p, double pK, double rK, double cK)
{
    if (fabs(p - pK) < EPS) return rK * cK;
    return k->gm * rK * cK * (1.0 - p / pK) / (1.0 - pow((p / pK), k->gm));
}

double o_godunov(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
                 double n1, double n2, o_d2* Ps, o_d2* Pu, double rad)
{
    const gconst K = make_gconst(gamma);
    const gconst* k = &K;
    gstate G;
    godunov_setup(k, Ls, Lu, Rs, Ru, n1, n2, &G);
    const double w = 0.0;
    const double enI = G.enI, pI = G.pI, rI = G.rI, enII = G.enII, pII = G.pII, rII = G.rII;
    const double cI = G.cI, cII = G.cII, teI2 = G.teI2, teII2 = G.teII2;
    double p = G.p, r = 0.0, te2 = 0.0, en = 0.0, time = 0.0, UU;

    if (G.il == 1 && G.ip == 1) {
        /* two shocks, kernel.cu:320-374 */
        double aI = sqrt(rI * (k->g2 / 2.0 * p + k->g1 / 2.0 * pI));
        double aII = sqrt(rII * (k->g2 / 2.0 * p + k->g1 / 2.0 * pII));
        double u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        double dI = enI - aI / rI;
        double dII = enII + aII / rII;
        UU = fmax(fabs(dI), fabs(dII));
        time = (UU > EPS8) ? krit * rad / UU : krit * rad / EPS8;
        if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; }
        else if ((w > dI) && (w <= u)) { en = u; r = rI * aI / (aI - rI * (enI - u)); te2 = teI2; }
        else if ((w > u) && (w < dII)) { en = u; r = rII * aII / (aII + rII * (enII - u)); te2 = teII2; }
        else if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; }
    } else if (G.il == 1 && G.ip == 0) {
        /* left shock, right fan, kernel.cu:378-450 */
        double aI = sqrt(rI * (k->g2 / 2.0 * p + k->g1 / 2.0 * pI));
        double aII = fan_a(k, p, pII, rII, cII);
        double u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        double dI = enI - aI / rI;
        double dII = enII + cII;
        double ddII = u + cII - k->g1 * (enII - u) / 2.0;
        UU = fmax(fabs(dI), fabs(dII));
        UU = fmax(UU, fabs(ddII));
        time = (UU > EPS8) ? krit * rad / UU : krit * rad / EPS8;
        if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; }
        if ((w > dI) && (w <= u)) { en = u; r = rI * aI / (aI - rI * (enI - u)); te2 = teI2; }
        if ((w > u) && (w <= ddII)) {
            double ce = cII - k->g1 / 2.0 * (enII - u);
            en = u; r = k->ga * p / ce / ce; te2 = teII2;
        }
        if ((w > ddII) && (w < dII)) {
            double ce = -k->g1 / k->g2 * (enII - w) + 2.0 / k->g2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / k->gm)); r = k->ga * p / ce / ce; te2 = teII2;
        }
        if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; }
    } else if (G.il == 0 && G.ip == 0) {
        /* two fans, kernel.cu:452-563 */
        double aI = fan_a(k, p, pI, rI, cI);
        double aII = fan_a(k, p, pII, rII, cII);
        double u = (aI * enI + aII * enII + pI - pII) / (aI + aII);
        double dI = enI - cI;
        double ddI = u - cI - k->g1 * (enI - u) / 2.0;
        double dII = enII + cII;
        double ddII = u + cII - k->g1 * (enII - u) / 2.0;
        UU = fmax(fabs(dI), fabs(dII));
        UU = fmax(UU, fabs(ddII));
        UU = fmax(UU, fabs(ddI));
        time = (UU > EPS8) ? krit * rad / UU : krit * rad / EPS8;
        if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; }
        if ((w > dI) && (w < ddI)) {
            double ce = k->g1 / k->g2 * (enI - w) + 2.0 / k->g2 * cI;
            en = w + ce; p = pI * pow((ce / cI), (1.0 / k->gm)); r = k->ga * p / ce / ce; te2 = teI2;
        }
        if ((w >= ddI) && (w <= u)) {
            double ce = cI + k->g1 / 2.0 * (enI - u);
            en = u; r = k->ga * p / ce / ce; te2 = teI2;
        }
        if ((w > u) && (w <= ddII)) {
            double ce = cII - k->g1 / 2.0 * (enII - u);
            en = u; r = k->ga * p / ce / ce; te2 = teII2;
        }
        if ((w > ddII) && (w < dII)) {
            double ce = -k->g1 / k->g2 * (enII - w) + 2.0 / k->g2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / k->gm)); r = k->ga * p / ce / ce; te2 = teII2;
        }
        if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; }
    } else if (G.il == -1 && G.ip == -1) {
        /* vacuum, kernel.cu:566-630 */
        double dI = enI - cI;
        double ddI = enI + 2.0 / k->g1 * cI;
        double dII = enII + cII;
        double ddII = enII - 2.0 / k->g1 * cII;
        UU = fmax(fabs(dI), fabs(dII));
        UU = fmax(UU, fabs(ddII));
        UU = fmax(UU, fabs(ddI));
        time = (UU > EPS8) ? krit * rad / UU : krit * rad / EPS8;
        if (w <= dI) { en = enI; p = pI; r = rI; te2 = teI2; }
        if ((w > dI) && (w < ddI)) {
            double ce = k->g1 / k->g2 * (enI - w) + 2.0 / k->g2 * cI;
            en = w + ce; p = pI * pow((ce / cI), (1.0 / k->gm)); r = k->ga * p / ce / ce; te2 = teI2;
        }
        if ((w >= ddI) && (w <= ddII)) { en = w; p = 0.0; r = 0.0; te2 = 0.0; }
        if ((w > ddII) && (w < dII)) {
            double ce = -k->g1 / k->g2 * (enII - w) + 2.0 / k->g2 * cII;
            en = w - ce; p = pII * pow((ce / cII), (1.0 / k->gm)); r = k->ga * p / ce / ce; te2 = teII2;
        }
        if (w >= dII) { en = enII; p = pII; r = rII; te2 = teII2; }
    } else {
        diag_inc(&g_diag.bad_branch);
        Ps->x = Ps->y = Pu->x = Pu->y = NAN;
        return NAN;
    }

    if (G.ipiz == 1) en = -en;   /* kernel.cu:633-637 */

    /* kernel.cu:639-651 (third frame vector and w-velocity are zero) */
    double uo = n1 * en + (-n2) * te2;
    double vo = n2 * en + n1 * te2;
    double eo = p / k->g1 + 0.5 * r * (uo * uo + vo * vo);
    en = n1 * uo + n2 * vo;
    Ps->x = (r * (en - w));
    Pu->x = (r * (en - w) * uo + n1 * p);
    Pu->y = (r * (en - w) * vo + n2 * p);
    Ps->y = ((en - w) * eo + en * p);
    return time;
}

void o_godunov_star(double gamma, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru, double n1, double n2,
                    double* pstar, double* ustar, int* pattern)
{
    const gconst K = make_gconst(gamma);
    gstate G;
    godunov_setup(&K, Ls, Lu, Rs, Ru, n1, n2, &G);
    double p = G.p, u = NAN;
    if (G.il == 1 && G.ip == 1) {
        double aI = sqrt(G.rI * (K.g2 / 2.0 * p + K.g1 / 2.0 * G.pI));
        double aII = sqrt(G.rII * (K.g2 / 2.0 * p + K.g1 / 2.0 * G.pII));
        u = (aI * G.enI + aII * G.enII + G.pI - G.pII) / (aI + aII);
    } else if (G.il == 1 && G.ip == 0) {
        double aI = sqrt(G.rI * (K.g2 / 2.0 * p + K.g1 / 2.0 * G.pI));
        double aII = fan_a(&K, p, G.pII, G.rII, G.cII);
        u = (aI * G.enI + aII * G.enII + G.pI - G.pII) / (aI + aII);
    } else if (G.il == 0 && G.ip == 0) {
        double aI = fan_a(&K, p, G.pI, G.rI, G.cI);
        double aII = fan_a(&K, p, G.pII, G.rII, G.cII);
        u = (aI * G.enI + aII * G.enII + G.pI - G.pII) / (aI + aII);
    }
    if (G.ipiz) u = -u;
    *pstar = p;
    *ustar = u;
    *pattern = (G.il == -1) ? 3 : (G.il == 0 ? 0 : (G.ip == 0 ? 1 : 2));
}

/* ------------------------------------------------------------------------ */
/* HLLC-MHD / HLLD: HLLDQ_Korolkov, Gas_Dinamic.cu:964-1510. The passive      */
/* scalar Q (always 1, its flux PQ discarded by every caller, kernel.cu:2061, */
/* 2550) is left out.                                                         */
/* ------------------------------------------------------------------------ */
double o_hlld(double gamma, double krit,
              double ro_L, double p_L, double v1_L, double v2_L, double v3_L,
              double Bx_L, double By_L, double Bz_L,
              double ro_R, double p_R, double v1_R, double v2_R, double v3_R,
              double Bx_R, double By_R, double Bz_R,
              double* P, double n1, double n2, double n3, double rad, int metod)
{
    const double g1 = gamma - 1.0;
    for (int i = 0; i < 8; ++i) P[i] = 0.0;
    if (metod != 2 && metod != 3) {
        /* the reference falls off the end of the function here (UB) */
        diag_inc(&g_diag.bad_branch);
        return -1.0;
    }
    /* Gas_Dinamic.cu:970-976 */
    double bx_L = Bx_L / SPI4, by_L = By_L / SPI4, bz_L = Bz_L / SPI4;
    double bx_R = Bx_R / SPI4, by_R = By_R / SPI4, bz_R = Bz_R / SPI4;

    /* Gas_Dinamic.cu:978-1019 */
    double t1 = 0, t2 = 0, t3 = 0, m1 = 0, m2 = 0, m3 = 0;
    if (n1 > 0.1) { t2 = 1.0; m3 = 1.0; }
    else if (n2 > 0.1) { t3 = 1.0; m1 = 1.0; }
    else if (n3 > 0.1) { t1 = 1.0; m2 = 1.0; }
    else if (n1 < -0.1) { t3 = -1.0; m2 = -1.0; }
    else if (n2 < -0.1) { t1 = -1.0; m3 = -1.0; }
    else if (n3 < -0.1) { t1 = -1.0; m2 = -1.0; }
    else { diag_inc(&g_diag.bad_branch); }

    /* Gas_Dinamic.cu:1022-1036 */
    double u1 = v1_L * n1 + v2_L * n2 + v3_L * n3;
    double v1 = v1_L * t1 + v2_L * t2 + v3_L * t3;
    double w1 = v1_L * m1 + v2_L * m2 + v3_L * m3;
    double u2 = v1_R * n1 + v2_R * n2 + v3_R * n3;
    double v2 = v1_R * t1 + v2_R * t2 + v3_R * t3;
    double w2 = v1_R * m1 + v2_R * m2 + v3_R * m3;
    double bn1 = bx_L * n1 + by_L * n2 + bz_L * n3;
    double bt1 = bx_L * t1 + by_L * t2 + bz_L * t3;
    double bm1 = bx_L * m1 + by_L * m2 + bz_L * m3;
    double bn2 = bx_R * n1 + by_R * n2 + bz_R * n3;
    double bt2 = bx_R * t1 + by_R * t2 + bz_R * t3;
    double bm2 = bx_R * m1 + by_R * m2 + bz_R * m3;

    /* Gas_Dinamic.cu:1047-1075 */
    double bb_L = bx_L * bx_L + by_L * by_L + bz_L * bz_L;
    double bb_R = bx_R * bx_R + by_R * by_R + bz_R * bz_R;
    double uu_L = (v1_L * v1_L + v2_L * v2_L + v3_L * v3_L) / 2.0;
    double uu_R = (v1_R * v1_R + v2_R * v2_R + v3_R * v3_R) / 2.0;
    double cfL = sqrt((gamma * p_L + bb_L +
                       sqrt((gamma * p_L + bb_L) * (gamma * p_L + bb_L) - 4.0 * gamma * p_L * (bn1 * bn1))) /
                      (2.0 * ro_L));
    double cfR = sqrt((gamma * p_R + bb_R +
                       sqrt((gamma * p_R + bb_R) * (gamma * p_R + bb_R) - 4.0 * gamma * p_R * (bn2 * bn2))) /
                      (2.0 * ro_R));
    double SL = fmin(u1, u2) - fmax(cfL, cfR);
    double SR = fmax(u1, u2) + fmax(cfL, cfR);
    double pTL = p_L + bb_L / 2.0;
    double pTR = p_R + bb_R / 2.0;
    double suR = (SR - u2);
    double suL = (SL - u1);
    double SM = (suR * ro_R * u2 - suL * ro_L * u1 - pTR + pTL) / (suR * ro_R - suL * ro_L);
    double PTT = (suR * ro_R * pTL - suL * ro_L * pTR + ro_L * ro_R * suR * suL * (u2 - u1)) /
                 (suR * ro_R - suL * ro_L);
    double UU = fmax(fabs(SL), fabs(SR));
    double time = krit * rad / UU;

    /* Gas_Dinamic.cu:1080-1125 (index 8 = Q dropped) */
    double FL[8], FR[8], UL[8], UR[8], PO[8];
    double e1 = p_L / g1 + ro_L * uu_L + bb_L / 2.0;
    double e2 = p_R / g1 + ro_R * uu_R + bb_R / 2.0;
    FL[0] = ro_L * u1;
    FL[1] = ro_L * u1 * u1 + pTL - bn1 * bn1;
    FL[2] = ro_L * u1 * v1 - bn1 * bt1;
    FL[3] = ro_L * u1 * w1 - bn1 * bm1;
    FL[4] = (e1 + pTL) * u1 - bn1 * (u1 * bn1 + v1 * bt1 + w1 * bm1);
    FL[5] = 0.0;
    FL[6] = u1 * bt1 - v1 * bn1;
    FL[7] = u1 * bm1 - w1 * bn1;
    FR[0] = ro_R * u2;
    FR[1] = ro_R * u2 * u2 + pTR - bn2 * bn2;
    FR[2] = ro_R * u2 * v2 - bn2 * bt2;
    FR[3] = ro_R * u2 * w2 - bn2 * bm2;
    FR[4] = (e2 + pTR) * u2 - bn2 * (u2 * bn2 + v2 * bt2 + w2 * bm2);
    FR[5] = 0.0;
    FR[6] = u2 * bt2 - v2 * bn2;
    FR[7] = u2 * bm2 - w2 * bn2;
    UL[0] = ro_L; UL[1] = ro_L * u1; UL[2] = ro_L * v1; UL[3] = ro_L * w1;
    UL[4] = e1; UL[5] = bn1; UL[6] = bt1; UL[7] = bm1;
    UR[0] = ro_R; UR[1] = ro_R * u2; UR[2] = ro_R * v2; UR[3] = ro_R * w2;
    UR[4] = e2; UR[5] = bn2; UR[6] = bt2; UR[7] = bm2;

    /* Gas_Dinamic.cu:1127-1135 */
    double bn = (SR * UR[5] - SL * UL[5] + FL[5] - FR[5]) / (SR - SL);
    double bt = (SR * UR[6] - SL * UL[6] + FL[6] - FR[6]) / (SR - SL);
    double bm = (SR * UR[7] - SL * UL[7] + FL[7] - FR[7]) / (SR - SL);
    double bbn = bn * bn;
    double ro_LL = ro_L * (SL - u1) / (SL - SM);
    double ro_RR = ro_R * (SR - u2) / (SR - SM);

    if (metod == 2) {
        /* HLLC-MHD with HLL-averaged tangential velocity, Gas_Dinamic.cu:1137-1301 */
        double sbv1 = u1 * bn1 + v1 * bt1 + w1 * bm1;
        double sbv2 = u2 * bn2 + v2 * bt2 + w2 * bm2;
        double UZ0 = (SR * UR[0] - SL * UL[0] + FL[0] - FR[0]) / (SR - SL);
        double UZ1 = (SR * UR[1] - SL * UL[1] + FL[1] - FR[1]) / (SR - SL);
        double UZ2 = (SR * UR[2] - SL * UL[2] + FL[2] - FR[2]) / (SR - SL);
        double UZ3 = (SR * UR[3] - SL * UL[3] + FL[3] - FR[3]) / (SR - SL);
        double suRm = suR / (SR - SM);
        double suLm = suL / (SL - SM);
        double ptzR = pTR + ro_R * suR * (SM - u2);
        double ptzL = pTL + ro_L * suL * (SM - u1);
        double ptz = (ptzR + ptzL) / 2.0;
        double vRR = UZ2 / UZ0, wRR = UZ3 / UZ0, vLL = vRR, wLL = wRR;
        double btt2 = bt, bmm2 = bm, btt1 = btt2, bmm1 = bmm2;
        double sbvz = (bn * UZ1 + bt * UZ2 + bm * UZ3) / UZ0;
        double ee2 = e2 * suRm + (ptz * SM - pTR * u2 + bn * (sbv2 - sbvz)) / (SR - SM);
        double ee1 = e1 * suLm + (ptz * SM - pTL * u1 + bn * (sbv1 - sbvz)) / (SL - SM);
        double ULL[8] = {ro_LL, ro_LL * SM, ro_LL * vLL, ro_LL * wLL, ee1, bn, btt1, bmm1};
        double URR[8] = {ro_RR, ro_RR * SM, ro_RR * vRR, ro_RR * wRR, ee2, bn, btt2, bmm2};
        if (SL >= 0.0) {
            for (int i = 0; i < 8; i++) PO[i] = FL[i];
        } else if (SL < 0.0 && SM >= 0.0) {
            for (int i = 0; i < 8; i++) PO[i] = FL[i] + SL * ULL[i] - SL * UL[i];
        } else if (SR > 0.0 && SM < 0.0) {
            for (int i = 0; i < 8; i++) PO[i] = FR[i] + SR * URR[i] - SR * UR[i];
        } else if (SR <= 0.0) {
            for (int i = 0; i < 8; i++) PO[i] = FR[i];
        } else {
            diag_inc(&g_diag.bad_branch);
            for (int i = 0; i < 8; i++) PO[i] = NAN;
        }
    } else {
        /* HLLD, Gas_Dinamic.cu:1303-1484 */
        double ttL = ro_L * suL * (SL - SM) - bbn;
        double ttR = ro_R * suR * (SR - SM) - bbn;
        double vLL, wLL, vRR, wRR, btt1, bmm1, btt2, bmm2;
        if (fabs(ttL) >= 0.00001) {
            vLL = v1 - bn * bt1 * (SM - u1) / ttL;
            wLL = w1 - bn * bm1 * (SM - u1) / ttL;
            btt1 = bt1 * (ro_L * suL * suL - bbn) / ttL;
            bmm1 = bm1 * (ro_L * suL * suL - bbn) / ttL;
        } else {
            vLL = v1; wLL = w1; btt1 = 0.0; bmm1 = 0.0;
        }
        if (fabs(ttR) >= 0.00001) {
            vRR = v2 - bn * bt2 * (SM - u2) / ttR;
            wRR = w2 - bn * bm2 * (SM - u2) / ttR;
            btt2 = bt2 * (ro_R * suR * suR - bbn) / ttR;
            bmm2 = bm2 * (ro_R * suR * suR - bbn) / ttR;
        } else {
            vRR = v2; wRR = w2; btt2 = 0.0; bmm2 = 0.0;
        }
        double eLL = (e1 * suL + PTT * SM - pTL * u1 +
                      bn * ((u1 * bn1 + v1 * bt1 + w1 * bm1) - (SM * bn + vLL * btt1 + wLL * bmm1))) /
                     (SL - SM);
        double eRR = (e2 * suR + PTT * SM - pTR * u2 +
                      bn * ((u2 * bn2 + v2 * bt2 + w2 * bm2) - (SM * bn + vRR * btt2 + wRR * bmm2))) /
                     (SR - SM);
        double sqrtroLL = sqrt(ro_LL);
        double sqrtroRR = sqrt(ro_RR);
        double SLL = SM - fabs(bn) / sqrtroLL;
        double SRR = SM + fabs(bn) / sqrtroRR;
        double idbn;
        if (fabs(bn) > 0.000001) {
            idbn = 1.0 * sgn(bn);
        } else {
            idbn = 0.0; SLL = SM; SRR = SM;
        }
        double vLLL = (sqrtroLL * vLL + sqrtroRR * vRR + idbn * (btt2 - btt1)) / (sqrtroLL + sqrtroRR);
        double wLLL = (sqrtroLL * wLL + sqrtroRR * wRR + idbn * (bmm2 - bmm1)) / (sqrtroLL + sqrtroRR);
        double bttt = (sqrtroLL * btt2 + sqrtroRR * btt1 + idbn * sqrtroLL * sqrtroRR * (vRR - vLL)) /
                      (sqrtroLL + sqrtroRR);
        double bmmm = (sqrtroLL * bmm2 + sqrtroRR * bmm1 + idbn * sqrtroLL * sqrtroRR * (wRR - wLL)) /
                      (sqrtroLL + sqrtroRR);
        double eLLL = eLL - idbn * sqrtroLL * ((SM * bn + vLL * btt1 + wLL * bmm1) -
                                               (SM * bn + vLLL * bttt + wLLL * bmmm));
        double eRRR = eRR + idbn * sqrtroRR * ((SM * bn + vRR * btt2 + wRR * bmm2) -
                                               (SM * bn + vLLL * bttt + wLLL * bmmm));
        double ULL[8] = {ro_LL, ro_LL * SM, ro_LL * vLL, ro_LL * wLL, eLL, bn, btt1, bmm1};
        double URR[8] = {ro_RR, ro_RR * SM, ro_RR * vRR, ro_RR * wRR, eRR, bn, btt2, bmm2};
        double ULLL[8] = {ro_LL, ro_LL * SM, ro_LL * vLLL, ro_LL * wLLL, eLLL, bn, bttt, bmmm};
        double URRR[8] = {ro_RR, ro_RR * SM, ro_RR * vLLL, ro_RR * wLLL, eRRR, bn, bttt, bmmm};
        if (SL >= 0.0) {
            for (int i = 0; i < 8; i++) PO[i] = FL[i];
        } else if (SL < 0.0 && SLL >= 0.0) {
            for (int i = 0; i < 8; i++) PO[i] = FL[i] + SL * ULL[i] - SL * UL[i];
        } else if (SLL <= 0.0 && SM >= 0.0) {
            for (int i = 0; i < 8; i++) PO[i] = FL[i] + SLL * ULLL[i] - (SLL - SL) * ULL[i] - SL * UL[i];
        } else if (SM < 0.0 && SRR > 0.0) {
            for (int i = 0; i < 8; i++) PO[i] = FR[i] + SRR * URRR[i] - (SRR - SR) * URR[i] - SR * UR[i];
        } else if (SR > 0.0 && SRR <= 0.0) {
            for (int i = 0; i < 8; i++) PO[i] = FR[i] + SR * URR[i] - SR * UR[i];
        } else if (SR <= 0.0) {
            for (int i = 0; i < 8; i++) PO[i] = FR[i];
        } else {
            diag_inc(&g_diag.bad_branch);
            for (int i = 0; i < 8; i++) PO[i] = NAN;
        }
    }
    /* Gas_Dinamic.cu:1281-1299 / 1488-1506 */
    double SN = fmax(fabs(SL), fabs(SR));
    PO[5] = -SN * (bn2 - bn1);
    double Q[8];
    Q[1] = n1 * PO[1] + t1 * PO[2] + m1 * PO[3];
    Q[2] = n2 * PO[1] + t2 * PO[2] + m2 * PO[3];
    Q[3] = n3 * PO[1] + t3 * PO[2] + m3 * PO[3];
    Q[5] = SPI4 * (n1 * PO[5] + t1 * PO[6] + m1 * PO[7]);
    Q[6] = SPI4 * (n2 * PO[5] + t2 * PO[6] + m2 * PO[7]);
    Q[7] = SPI4 * (n3 * PO[5] + t3 * PO[6] + m3 * PO[7]);
    P[0] = PO[0];
    P[1] = Q[1];
    P[2] = Q[2];
    P[3] = Q[3];
    P[4] = Q[5];
    P[5] = Q[6];
    P[6] = Q[7];
    P[7] = PO[4];
    return time;
}

/* ------------------------------------------------------------------------ */
/* stencil gathering with boundary kinds                                      */
/* ------------------------------------------------------------------------ */

/* Map logical index k (may be -2..n+1) to a stored index; *flip = 1 when the
 * reflective rule applies (normal velocity / normal B change sign). The
 * reference's transmissive sides copy the edge cell for both ghost layers
 * (kernel.cu:1762-1793, 3374-3417); its symmetry sides mirror (kernel.cu:
 * 2263-2277, 2356-2405). Periodic sides wrap (no reference counterpart). */
/* Source index of position k on a line of n cells, and the kind of side it
 * came through (*side = the side's bc kind, or -1 inside the line): mirrored
 * for the reflective / symmetry / axis kinds, wrapped for periodic, the
 * edge cell otherwise. */
static inline int map_idx(int k, int n, int lo, int hi, int* side)
{
    *side = -1;
    if (k < 0) {
        if (lo == O_BC_PERIODIC) return ((k % n) + n) % n;
        *side = lo;
        if (lo == O_BC_REFLECTIVE || lo == O_BC_SYMMETRY || lo == O_BC_AXIS) return (-1 - k < n) ? -1 - k : n - 1;
        return 0;
    }
    if (k >= n) {
        if (hi == O_BC_PERIODIC) return k % n;
        *side = hi;
        if (hi == O_BC_REFLECTIVE || hi == O_BC_SYMMETRY || hi == O_BC_AXIS) return (2 * n - 1 - k >= 0) ? 2 * n - 1 - k : 0;
        return n - 1;
    }
    return k;
}

/* primitive cell: rho, p, u, v, Bx, By, Bz (Gaussian B) */
typedef struct { double r, p, u, v, bx, by, bz; } prim;

typedef struct {
    const o_grid* g;
    int kind;            /* 0 = reference AoS primitives, 1 = SoA conserved */
    int mhd;
    const o_d2 *s, *u;
    const o_d3* b;
    const double* U;
    int inflow_b_zero;   /* INFLOW ghosts: B = 0 (Kernel_TVD) instead of the edge cell's (add2) */
} src_t;

static inline prim cons_cell_to_prim(const src_t* S, long k)
{
    const o_grid* g = S->g;
    const long plane = (long)g->nx * g->ny;
    prim q;
    const double* U = S->U;
    double r = U[k], mx = U[plane + k], my = U[2 * plane + k];
    double inv = 1.0 / r;
    q.r = r;
    q.u = mx * inv;
    q.v = my * inv;
    if (S->mhd) {
        q.bx = U[3 * plane + k]; q.by = U[4 * plane + k]; q.bz = U[5 * plane + k];
        double E = U[6 * plane + k];
        double ke = 0.5 * (mx * q.u + my * q.v);
        double me = (q.bx * q.bx + q.by * q.by + q.bz * q.bz) / CPI8;
        q.p = (g->gamma - 1.0) * (E - ke - me);
    } else {
        q.bx = q.by = q.bz = 0.0;
        double E = U[3 * plane + k];
        double ke = 0.5 * (mx * q.u + my * q.v);
        q.p = (g->gamma - 1.0) * (E - ke);
    }
    return q;
}

static inline prim fetch(const src_t* S, int i, int j)
{
    const o_grid* g = S->g;
    int sx, sy;
    int ii = map_idx(i, g->nx, g->bc[0], g->bc[1], &sx);
    int jj = map_idx(j, g->ny, g->bc[2], g->bc[3], &sy);
    long k = (long)jj * g->nx + ii;
    prim q;
    if (S->kind == 0) {
        q.r = S->s[k].x; q.p = S->s[k].y; q.u = S->u[k].x; q.v = S->u[k].y;
        if (S->mhd) { q.bx = S->b[k].x; q.by = S->b[k].y; q.bz = S->b[k].z; }
        else { q.bx = q.by = q.bz = 0.0; }
    } else {
        q = cons_cell_to_prim(S, k);
    }
    /* ghost rules by side kind (kernel.cu:1960-2032, 2232-2277, 2322-2405) */
    if (sx == O_BC_REFLECTIVE) { q.u = -q.u; q.bx = -q.bx; }
    else if (sx == O_BC_SYMMETRY) { q.u = -q.u; q.bx = -q.bx; q.by = -q.by; }
    else if (sx == O_BC_OUTFLOW) { if (q.u < 0) q.u = 0.01; }
    if (sy == O_BC_REFLECTIVE) { q.v = -q.v; q.by = -q.by; }
    else if (sy == O_BC_AXIS) { q.v = -q.v; q.bz = -q.bz; }
    else if (sy == O_BC_INFLOW) {
        q.r = 1.0; q.p = 1.0 / g->gamma; q.u = 0.0; q.v = 0.0;
        if (S->inflow_b_zero) { q.bx = 0.0; q.by = 0.0; q.bz = 0.0; }
    }
    return q;
}

/* plus-shaped radius-2 stencil: c, west 1/2, east 1/2, south 1/2, north 1/2 */
typedef struct { prim c, w1, w2, e1, e2, s1, s2, n1, n2; } stencil;

static void gather(const src_t* S, int i, int j, int order, stencil* st)
{
    st->c = fetch(S, i, j);
    st->e1 = fetch(S, i + 1, j);
    st->w1 = fetch(S, i - 1, j);
    st->s1 = fetch(S, i, j - 1);
    st->n1 = fetch(S, i, j + 1);
    if (order >= 2) {
        st->e2 = fetch(S, i + 2, j);
        st->w2 = fetch(S, i - 2, j);
        st->s2 = fetch(S, i, j - 2);
        st->n2 = fetch(S, i, j + 2);
    }
}

/* Face states for the four faces: own-side (f12 east, f13 south, f14 west,
 * f15 north) and neighbour-side (f21, f31, f41, f51). Order 2 restates the
 * minmod MUSCL block kernel.cu:3436-3544 (Euler) / 2407-2548 (MHD) with
 * absolute coordinates and the rho/p <= 0 fallback (grad_p, kernel.cu:28). */
typedef struct { prim f12, f13, f14, f15, f21, f31, f41, f51; } faces;

static void reconstruct(const o_grid* g, int i, int j, int order, int mhd, const stencil* st, faces* F)
{
    if (order < 2) {
        F->f12 = F->f13 = F->f14 = F->f15 = st->c;
        F->f21 = st->e1; F->f31 = st->s1; F->f41 = st->w1; F->f51 = st->n1;
        return;
    }
    const double dx = g->dx, dy = g->dy;
    const double x = g->x0 + i * dx;
    const double y = g->y0 + j * dy;
    double A, B;
    const prim *c = &st->c;
    /* x: own faces */
#define L2X(fld, dst12, dst14) \
    o_linear2(x - dx, st->w1.fld, x, c->fld, x + dx, st->e1.fld, x - dx / 2.0, x + dx / 2.0, &A, &B); \
    F->dst12.fld = B; F->dst14.fld = A;
#define L2Y(fld) \
    o_linear2(y - dy, st->s1.fld, y, c->fld, y + dy, st->n1.fld, y - dy / 2.0, y + dy / 2.0, &A, &B); \
    F->f15.fld = B; F->f13.fld = A;
    L2X(r, f12, f14);
    if (F->f12.r <= 0) F->f12.r = c->r;
    if (F->f14.r <= 0) F->f14.r = c->r;
    L2X(p, f12, f14);
    if (F->f12.p <= 0) F->f12.p = c->p;
    if (F->f14.p <= 0) F->f14.p = c->p;
    L2X(u, f12, f14);
    L2X(v, f12, f14);
    L2Y(r);
    if (F->f15.r <= 0) F->f15.r = c->r;
    if (F->f13.r <= 0) F->f13.r = c->r;
    L2Y(p);
    if (F->f15.p <= 0) F->f15.p = c->p;
    if (F->f13.p <= 0) F->f13.p = c->p;
    L2Y(u);
    L2Y(v);
    if (mhd) {
        L2X(bx, f12, f14); L2X(by, f12, f14); L2X(bz, f12, f14);
        L2Y(bx); L2Y(by); L2Y(bz);
    } else {
        F->f12.bx = F->f12.by = F->f12.bz = 0.0;
        F->f13.bx = F->f13.by = F->f13.bz = 0.0;
        F->f14.bx = F->f14.by = F->f14.bz = 0.0;
        F->f15.bx = F->f15.by = F->f15.bz = 0.0;
    }
#undef L2X
#undef L2Y
    /* neighbour faces (linear): kernel.cu:3518-3544 */
#define NB(dst, nb1, nb2, X, D, fld) \
    F->dst.fld = o_linear(X, c->fld, X + (D), st->nb1.fld, X + 2.0 * (D), st->nb2.fld, X + (D) / 2.0);
#define NBALL(dst, nb1, nb2, X, D) \
    NB(dst, nb1, nb2, X, D, r) if (F->dst.r <= 0) F->dst.r = st->nb1.r; \
    NB(dst, nb1, nb2, X, D, p) if (F->dst.p <= 0) F->dst.p = st->nb1.p; \
    NB(dst, nb1, nb2, X, D, u) NB(dst, nb1, nb2, X, D, v) \
    if (mhd) { NB(dst, nb1, nb2, X, D, bx) NB(dst, nb1, nb2, X, D, by) NB(dst, nb1, nb2, X, D, bz) } \
    else { F->dst.bx = F->dst.by = F->dst.bz = 0.0; }
    NBALL(f21, e1, e2, x, dx)
    NBALL(f41, w1, w2, x, -dx)
    NBALL(f31, s1, s2, y, -dy)
    NBALL(f51, n1, n2, y, dy)
#undef NB
#undef NBALL
}

/* Euler flux along an outward normal; returns the face time step. */
static inline double euler_flux(const o_grid* g, int method, const prim* L, const prim* R,
                                double n1, double n2, o_d2* Ps, o_d2* Pu, double rad)
{
    o_d2 Ls = {L->r, L->p}, Lu = {L->u, L->v}, Rs = {R->r, R->p}, Ru = {R->u, R->v};
    if (method == 2) return o_godunov(g->gamma, g->krit, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
    if (method == 0) return o_hll(g->gamma, g->krit, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
    /* the HLLC variants as further methods (no reference kernel calls them) */
    if (method == 3) return o_hllc_korolkov(g->gamma, g->krit, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
    if (method == 4) return o_hllcq_korolkov_2d(g->gamma, g->krit, Ls, Lu, Rs, Ru, 0.0, 0.0, n1, n2, Ps, Pu, NULL, rad);
    return o_hllc(g->gamma, g->krit, Ls, Lu, Rs, Ru, n1, n2, Ps, Pu, rad);
}

typedef struct { o_d2 PS, PU; double tmin; } euler_sum;

/* four outward-normal fluxes of one cell, summed with face lengths in the
 * reference's order (kernel.cu:1796-1857, 3547-3647). radswap = 1 passes dy
 * on x-faces and dx on y-faces as Ker_Dekard / Kernel_TVD_old do. */
static euler_sum euler_cell_fluxes(const o_grid* g, int method, const faces* F, int radswap)
{
    const double dx = g->dx, dy = g->dy;
    const double rx = radswap ? dy : dx, ry = radswap ? dx : dy;
    o_d2 Ps12 = {0, 0}, Pu12 = {0, 0}, Ps13 = {0, 0}, Pu13 = {0, 0};
    o_d2 Ps14 = {0, 0}, Pu14 = {0, 0}, Ps15 = {0, 0}, Pu15 = {0, 0};
    double tmin = 1000;
    tmin = fmin(tmin, euler_flux(g, method, &F->f12, &F->f21, 1, 0, &Ps12, &Pu12, rx));
    tmin = fmin(tmin, euler_flux(g, method, &F->f13, &F->f31, 0, -1, &Ps13, &Pu13, ry));
    tmin = fmin(tmin, euler_flux(g, method, &F->f14, &F->f41, -1, 0, &Ps14, &Pu14, rx));
    tmin = fmin(tmin, euler_flux(g, method, &F->f15, &F->f51, 0, 1, &Ps15, &Pu15, ry));
    euler_sum r;
    r.PS.x = Ps12.x * dy + Ps13.x * dx + Ps14.x * dy + Ps15.x * dx;
    r.PS.y = Ps12.y * dy + Ps13.y * dx + Ps14.y * dy + Ps15.y * dx;
    r.PU.x = Pu12.x * dy + Pu13.x * dx + Pu14.x * dy + Pu15.x * dx;
    r.PU.y = Pu12.y * dy + Pu13.y * dx + Pu14.y * dy + Pu15.y * dx;
    r.tmin = tmin;
    return r;
}

/* The axisymmetric steps divide by the signed cell centre y0 + m dy; no row
 * may be centred on the axis (the reference would divide by zero there). */
static int axis_rows_ok(const o_grid* g)
{
    if (!isfinite(g->y0)) return 0;
    const double k = nearbyint(-g->y0 / g->dy);
    return !(k >= 0.0 && k < (double)g->ny && fabs(g->y0 + k * g->dy) <= 1e-9 * g->dy);
}

static int is_frozen(const o_grid* g, int i, int j)
{
    return (g->bc[0] == O_BC_FROZEN && i == 0) || (g->bc[1] == O_BC_FROZEN && i == g->nx - 1) ||
           (g->bc[2] == O_BC_FROZEN && j == 0) || (g->bc[3] == O_BC_FROZEN && j == g->ny - 1);
}

/* Kernel_TVD_old's subsonic rule for its west ghost (kernel.cu:3332-3336,
 * 3372-3376): vx = -5 where the cell is subsonic and vx > -5. */
static void west_subsonic_rule(const o_grid* g, prim* q)
{
    double Max = sqrt((q->u * q->u + q->v * q->v) / (g->gamma * q->p / q->r));
    if ((q->u > -5.0) && (Max < 1)) q->u = -5.0;
}

/* Kernel_TVD_old's west side (kernel.cu:3323-3337, 3365-3378): the cell of
 * column 0 sees both ghost columns as its own density and vy, the pressure
 * 1/(ggg M_inf^2) and its vx under the subsonic rule; the cell of column 1
 * sees column -1 as the cell of column 0 under that rule, pressure kept. */
static void west_inflow(const o_grid* g, int i, stencil* st)
{
    if (i == 0) {
        prim q = st->c;
        west_subsonic_rule(g, &q);
        q.p = 1.0 / (g->gamma * g->m_inf * g->m_inf);
        st->w1 = q;
        st->w2 = q;
    } else if (i == 1) {
        prim q = st->w1;
        west_subsonic_rule(g, &q);
        st->w2 = q;
    }
}

/* Euler reference-layout step shared by Ker_Dekard (order 1) and
 * Kernel_TVD_old (order 2). Frozen edge cells (the reference's "hard"
 * boundary rows, kernel.cu:1752-1758) are copied. Kernel_TVD_old's inner
 * sphere (kernel.cu:3303) and axisymmetric 1/y source terms
 * (kernel.cu:3651-3662) apply when the grid asks for them (order 2 only). */
static int euler_ref_step(const o_grid* g, int order, const o_d2* s, const o_d2* u, o_d2* s2, o_d2* u2,
                          double* T, const double* T_do, int method)
{
    if (method < 0 || method > 4) return -1;
    if ((g->axisym || g->r_frozen != 0.0) && order != 2) return -1;
    if (g->axisym && !axis_rows_ok(g)) return -1;
    const int west_in = g->bc[0] == O_BC_WEST_INFLOW;
    if (west_in && (order != 2 || !(g->m_inf > 0.0) || !isfinite(g->m_inf))) return -1;
    for (int k = 1; k < 4; ++k)
        if (g->bc[k] == O_BC_WEST_INFLOW) return -1;
    src_t S = {g, 0, 0, s, u, NULL, NULL, 0};
    const double dx = g->dx, dy = g->dy, gam = g->gamma;
    const double Tdo = *T_do;
    double tglob = *T;
#pragma omp parallel for schedule(static) reduction(min : tglob)
    for (int j = 0; j < g->ny; ++j) {
        for (int i = 0; i < g->nx; ++i) {
            long k = (long)j * g->nx + i;
            const double x = g->x0 + i * dx, y = g->y0 + j * dy;
            const int sphere = g->r_frozen > 0.0 && sqrt(x * x + y * y) < g->r_frozen;
            if (is_frozen(g, i, j) || sphere) { s2[k] = s[k]; u2[k] = u[k]; continue; }
            stencil st;
            faces F;
            gather(&S, i, j, order, &st);
            if (west_in) west_inflow(g, i, &st);
            reconstruct(g, i, j, order, 0, &st, &F);
            euler_sum E = euler_cell_fluxes(g, method, &F, 1);
            if (tglob > E.tmin) tglob = E.tmin;   /* kernel.cu:1840-1844 (exact min) */
            /* kernel.cu:1859-1874; with the 1/y terms kernel.cu:3651-3662 */
            double dV = dx * dy;
            double r2, ux, uy, p2;
            if (g->axisym) {
                r2 = s[k].x - (Tdo / dV) * E.PS.x - Tdo * s[k].x * u[k].y / y;
                if (r2 <= 0) { diag_inc(&g_diag.clamp_rho); r2 = 0.0001; }
                ux = (s[k].x * u[k].x - (Tdo / dV) * E.PU.x - Tdo * s[k].x * u[k].y * u[k].x / y) / r2;
                uy = (s[k].x * u[k].y - (Tdo / dV) * E.PU.y - Tdo * s[k].x * u[k].y * u[k].y / y) / r2;
                p2 = (((s[k].y / (gam - 1) + s[k].x * (u[k].x * u[k].x + u[k].y * u[k].y) * 0.5) -
                       (Tdo / dV) * E.PS.y -
                       Tdo * u[k].y * (gam * s[k].y / (gam - 1) + s[k].x * (u[k].x * u[k].x + u[k].y * u[k].y) * 0.5) /
                           y) -
                      0.5 * r2 * (ux * ux + uy * uy)) * (gam - 1);
            } else {
                r2 = s[k].x - (Tdo / dV) * E.PS.x;
                if (r2 <= 0) { diag_inc(&g_diag.clamp_rho); r2 = 0.0001; }
                ux = (s[k].x * u[k].x - (Tdo / dV) * E.PU.x) / r2;
                uy = (s[k].x * u[k].y - (Tdo / dV) * E.PU.y) / r2;
                p2 = (((s[k].y / (gam - 1) + s[k].x * (u[k].x * u[k].x + u[k].y * u[k].y) * 0.5) -
                       (Tdo / dV) * E.PS.y) -
                      0.5 * r2 * (ux * ux + uy * uy)) * (gam - 1);
            }
            if (p2 <= 0) { diag_inc(&g_diag.clamp_p); p2 = 0.000001; }
            s2[k].x = r2; s2[k].y = p2; u2[k].x = ux; u2[k].y = uy;
        }
    }
    *T = tglob;
    return 0;
}

int o_ker_dekard(const o_grid* g, const o_d2* s, const o_d2* u, o_d2* s2, o_d2* u2,
                 double* T, const double* T_do, int method)
{
    return euler_ref_step(g, 1, s, u, s2, u2, T, T_do, method);
}

int o_kernel_tvd_old(const o_grid* g, const o_d2* s, const o_d2* u, o_d2* s2, o_d2* u2,
                     double* T, const double* T_do, int method)
{
    return euler_ref_step(g, 2, s, u, s2, u2, T, T_do, method);
}

/* MHD: four outward-normal fluxes + Powell divergence term, kernel.cu:2559-2604
 * (MUSCL) / 2069-2125 (first order). B-average uses the cell-centred
 * neighbours in both. */
typedef struct { o_d2 PS, PU; o_d3 PB; double Pdiv, tmin; } mhd_sum;

static inline double mhd_face(const o_grid* g, int metod, const prim* L, const prim* R, double* P,
                              double n1, double n2, double rad)
{
    return o_hlld(g->gamma, g->krit, L->r, L->p, L->u, L->v, 0.0, L->bx, L->by, L->bz,
                  R->r, R->p, R->u, R->v, 0.0, R->bx, R->by, R->bz, P, n1, n2, 0.0, rad, metod);
}

static mhd_sum mhd_cell_fluxes(const o_grid* g, int metod, const faces* F, const stencil* st)
{
    const double dx = g->dx, dy = g->dy;
    double P[8];
    mhd_sum m;
    m.PS.x = m.PS.y = m.PU.x = m.PU.y = 0.0;
    m.PB.x = m.PB.y = m.PB.z = 0.0;
    m.Pdiv = 0.0;
    m.tmin = 1000;
#define ACC(len) \
    m.PS.x = m.PS.x + P[0] * len; m.PS.y = m.PS.y + P[7] * len; \
    m.PU.x = m.PU.x + P[1] * len; m.PU.y = m.PU.y + P[2] * len; \
    m.PB.x = m.PB.x + P[4] * len; m.PB.y = m.PB.y + P[5] * len; m.PB.z = m.PB.z + P[6] * len;
    m.tmin = fmin(m.tmin, mhd_face(g, metod, &F->f12, &F->f21, P, 1.0, 0.0, dx));
    ACC(dy)
    m.Pdiv = m.Pdiv + dy * (st->c.bx + st->e1.bx) / 2.0;
    m.tmin = fmin(m.tmin, mhd_face(g, metod, &F->f13, &F->f31, P, 0.0, -1.0, dy));
    ACC(dx)
    m.Pdiv = m.Pdiv - dx * (st->c.by + st->s1.by) / 2.0;
    m.tmin = fmin(m.tmin, mhd_face(g, metod, &F->f14, &F->f41, P, -1.0, 0.0, dx));
    ACC(dy)
    m.Pdiv = m.Pdiv - dy * (st->c.bx + st->w1.bx) / 2.0;
    m.tmin = fmin(m.tmin, mhd_face(g, metod, &F->f15, &F->f51, P, 0.0, 1.0, dy));
    ACC(dx)
    m.Pdiv = m.Pdiv + dx * (st->c.by + st->n1.by) / 2.0;
#undef ACC
    return m;
}

/* add2's rows below the axis (kernel.cu:1885-1889, 2026-2057): the cell is
 * evaluated in the mirrored frame, its south and north neighbours swapped
 * and vy, By, Bz of the five cells negated; the update then takes |y| and
 * negates vy, By, Bz of the result (kernel.cu:2177-2182). */
static void mirror_stencil(stencil* st)
{
    prim t = st->s1;
    st->s1 = st->n1;
    st->n1 = t;
    prim* q[5] = {&st->c, &st->e1, &st->w1, &st->s1, &st->n1};
    for (int k = 0; k < 5; ++k) { q[k]->v = -q[k]->v; q[k]->by = -q[k]->by; q[k]->bz = -q[k]->bz; }
}

/* MHD reference-layout step: Cartesian form of add2 / Kernel_TVD update
 * (kernel.cu:2141-2195, 2619-2660 without the 1/y source terms). */
static int mhd_ref_step(const o_grid* g, int order, const o_d2* s, const o_d2* u, const o_d3* b,
                        o_d2* s2, o_d2* u2, o_d3* b2, double* T, const double* T_do, int method)
{
    if (method != 2 && method != 3) return -1;
    if (g->axisym && !axis_rows_ok(g)) return -1;
    for (int k = 0; k < 4; ++k)
        if (g->bc[k] == O_BC_WEST_INFLOW) return -1;
    src_t S = {g, 0, 1, s, u, b, NULL, order == 2};
    const double dx = g->dx, dy = g->dy, gam = g->gamma;
    const double Tdo = *T_do;
    double tglob = *T;
#pragma omp parallel for schedule(static) reduction(min : tglob)
    for (int j = 0; j < g->ny; ++j) {
        for (int i = 0; i < g->nx; ++i) {
            long k = (long)j * g->nx + i;
            const double x = g->x0 + i * dx, y = g->y0 + j * dy, y_signed = y;
            const int sphere = g->r_frozen > 0.0 && sqrt(x * x + y * y) < g->r_frozen;  /* kernel.cu:1919 */
            if (is_frozen(g, i, j) || sphere) { s2[k] = s[k]; u2[k] = u[k]; b2[k] = b[k]; continue; }
            stencil st;
            faces F;
            gather(&S, i, j, order, &st);
            /* add2 only: Kernel_TVD divides by the signed y (kernel.cu:2203) */
            const int mirror = order == 1 && g->axisym && y < 0.0;
            if (mirror) mirror_stencil(&st);
            reconstruct(g, i, j, order, 1, &st, &F);
            mhd_sum m = mhd_cell_fluxes(g, method, &F, &st);
            if (tglob > m.tmin) tglob = m.tmin;
            const prim* c = &st.c;
            double dV = dx * dy;
            /* U8 (Header.h:16), skk (Header.h:14), kvv (Header.h:15) with w = 0 */
            double U8 = c->p / (gam - 1.0) + 0.5 * c->r * (c->u * c->u + c->v * c->v + 0.0 * 0.0) +
                        (c->bx * c->bx + c->by * c->by + c->bz * c->bz) / 25.13274122871834590768;
            double skk = c->u * c->bx + c->v * c->by + 0.0 * c->bz;
            double r2, ux, uy, bx, by, bz, p2;
            if (g->axisym) {
                /* axisymmetric sources, kernel.cu:2134-2190 */
                const double y = mirror ? fabs(y_signed) : y_signed;
                const double Pdiv = m.Pdiv + dV * c->by / y;
                r2 = c->r - Tdo * m.PS.x / dV - Tdo * c->r * c->v / y;
                if (r2 <= 0) { diag_inc(&g_diag.clamp_rho); r2 = c->r; }
                ux = (c->r * c->u - Tdo * (m.PU.x + (c->bx / CPI4) * Pdiv) / dV -
                      Tdo * (c->r * c->u * c->v - c->bx * c->by / CPI4) / y) / r2;
                uy = (c->r * c->v - Tdo * (m.PU.y + (c->by / CPI4) * Pdiv) / dV -
                      Tdo * (c->r * c->v * c->v + (c->bz * c->bz - c->by * c->by) / CPI4) / y) / r2;
                bx = (c->bx - Tdo * (m.PB.x + c->u * Pdiv) / dV - Tdo * (c->v * c->bx - c->by * c->u) / y);
                by = (c->by - Tdo * (m.PB.y + c->v * Pdiv) / dV);
                bz = (c->bz - Tdo * (m.PB.z) / dV);
                p2 = (U8 - Tdo * (m.PS.y + (skk / CPI4) * Pdiv) / dV -
                      Tdo * (((U8 + c->p + (c->bx * c->bx + c->by * c->by + c->bz * c->bz) / CPI8) * c->v -
                              c->by * skk / CPI4) / y) -
                      0.5 * r2 * (ux * ux + uy * uy + 0.0 * 0.0) - (bx * bx + by * by + bz * bz) / CPI8) *
                     (gam - 1.0);
            } else {
                r2 = c->r - Tdo * m.PS.x / dV;
                if (r2 <= 0) { diag_inc(&g_diag.clamp_rho); r2 = c->r; }
                ux = (c->r * c->u - Tdo * (m.PU.x + (c->bx / CPI4) * m.Pdiv) / dV) / r2;
                uy = (c->r * c->v - Tdo * (m.PU.y + (c->by / CPI4) * m.Pdiv) / dV) / r2;
                bx = (c->bx - Tdo * (m.PB.x + c->u * m.Pdiv) / dV);
                by = (c->by - Tdo * (m.PB.y + c->v * m.Pdiv) / dV);
                bz = (c->bz - Tdo * (m.PB.z) / dV);
                p2 = (U8 - Tdo * (m.PS.y + (skk / CPI4) * m.Pdiv) / dV -
                      0.5 * r2 * (ux * ux + uy * uy + 0.0 * 0.0) - (bx * bx + by * by + bz * bz) / CPI8) *
                     (gam - 1.0);
            }
            if (mirror) { uy = -uy; by = -by; bz = -bz; }
            if (p2 <= 0) { diag_inc(&g_diag.clamp_p); p2 = 0.000001; }
            s2[k].x = r2; s2[k].y = p2; u2[k].x = ux; u2[k].y = uy;
            b2[k].x = bx; b2[k].y = by; b2[k].z = bz;
        }
    }
    *T = tglob;
    return 0;
}

int o_add2(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b,
           o_d2* s2, o_d2* u2, o_d3* b2, double* T, const double* T_do, int i, int method)
{
    (void)i;
    return mhd_ref_step(g, 1, s, u, b, s2, u2, b2, T, T_do, method);
}

int o_kernel_tvd(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b,
                 o_d2* s2, o_d2* u2, o_d3* b2, double* T, const double* T_do, int i, int method)
{
    (void)i;
    return mhd_ref_step(g, 2, s, u, b, s2, u2, b2, T, T_do, method);
}

/* ------------------------------------------------------------------------ */
/* Godunov-Kolgan-Rodionov predictor-corrector step: Kernel_GKR,            */
/* kernel.cu:2663-3284, with predictor (kernel.cu:1611-1720) and            */
/* POTOK_Korolkov (Gas_Dinamic.cu:70-181).                                  */
/* State vector q[7] = (rho, p, vx, vy, Bx, By, Bz), Gaussian B.            */
/* Choices where the source does not define the behaviour:                  */
/*  * POTOK_Korolkov fills P[0..7] (energy flux at P[7] after its SWAP,      */
/*    Gas_Dinamic.cu:175-179) while predictor declares double P[7]           */
/*    (kernel.cu:1613) and reads P[7] as the energy flux: P has 8 entries.   */
/*  * cell centres at x0 + n dx, y0 + m dy (the reference's Kernel_GKR puts  */
/*    x at x_min + n (x_max - x_min)/(N - 1)); the frozen radius is the      */
/*    grid's r_frozen (2.0 in Kernel_GKR, kernel.cu:2688).                   */
/* Kept as written: the vy slope is the pressure slope (DX*[3], DY*[3] from  */
/* s.y, kernel.cu:2965-3038); the second-neighbour ghosts of kernel.cu:      */
/* 2909-2957 (n = 1 takes cell n+1 mirrored, m = 0 takes the cell itself);   */
/* the predictor's and the corrector's 1/y source terms (always on), with    */
/* the ghost rows' predictors at y -/+ dy; the corrector starts from the     */
/* mean of the cell and its predictor and takes dt/2 (kernel.cu:3248-3276);  */
/* Powell's term computed and then zeroed (kernel.cu:3245).                  */
/* ------------------------------------------------------------------------ */
#define GKR_KV(x) ((x) * (x))
#define GKR_KVV(x, y, z) (GKR_KV(x) + GKR_KV(y) + GKR_KV(z))
#define GKR_SKK(u, v, w, bx, by, bz) ((u) * (bx) + (v) * (by) + (w) * (bz))
static inline double gkr_u8(double gam, double ro, double p, double u, double v, double w, double bx, double by,
                            double bz)
{
    /* Header.h:16 */
    return p / (gam - 1.0) + 0.5 * ro * GKR_KVV(u, v, w) + GKR_KVV(bx, by, bz) / CPI8;
}

/* POTOK_Korolkov (Gas_Dinamic.cu:70-181) with Q_L = 0: the physical MHD
 * flux of one state along an axis normal, P[8] = (F_rho, F_mx, F_my, F_mz,
 * F_Bx, F_By, F_Bz, F_E). */
void o_potok(double gam, double ro_L, double p_L, double v1_L, double v2_L, double v3_L, double Bx_L,
             double By_L, double Bz_L, double* P, double n1, double n2, double n3)
{
    const double g1 = gam - 1.0;
    double bx_L = Bx_L / SPI4, by_L = By_L / SPI4, bz_L = Bz_L / SPI4;
    double t1 = 0, t2 = 0, t3 = 0, m1 = 0, m2 = 0, m3 = 0;
    if (n1 > 0.1) { t2 = 1.0; m3 = 1.0; }
    else if (n2 > 0.1) { t3 = 1.0; m1 = 1.0; }
    else if (n3 > 0.1) { t1 = 1.0; m2 = 1.0; }
    else if (n1 < -0.1) { t3 = -1.0; m2 = -1.0; }
    else if (n2 < -0.1) { t1 = -1.0; m3 = -1.0; }
    else if (n3 < -0.1) { t1 = -1.0; m2 = -1.0; }
    else diag_inc(&g_diag.bad_branch);
    double u1 = v1_L * n1 + v2_L * n2 + v3_L * n3;
    double v1 = v1_L * t1 + v2_L * t2 + v3_L * t3;
    double w1 = v1_L * m1 + v2_L * m2 + v3_L * m3;
    double bn1 = bx_L * n1 + by_L * n2 + bz_L * n3;
    double bt1 = bx_L * t1 + by_L * t2 + bz_L * t3;
    double bm1 = bx_L * m1 + by_L * m2 + bz_L * m3;
    double bb_L = GKR_KV(bx_L) + GKR_KV(by_L) + GKR_KV(bz_L);
    double uu_L = (GKR_KV(v1_L) + GKR_KV(v2_L) + GKR_KV(v3_L)) / 2.0;
    double pTL = p_L + bb_L / 2.0;
    double e1 = p_L / g1 + ro_L * uu_L + bb_L / 2.0;
    double PO[8];
    PO[0] = ro_L * u1;
    PO[1] = ro_L * u1 * u1 + pTL - GKR_KV(bn1);
    PO[2] = ro_L * u1 * v1 - bn1 * bt1;
    PO[3] = ro_L * u1 * w1 - bn1 * bm1;
    PO[4] = (e1 + pTL) * u1 - bn1 * (u1 * bn1 + v1 * bt1 + w1 * bm1);
    PO[5] = 0.0;
    PO[6] = u1 * bt1 - v1 * bn1;
    PO[7] = u1 * bm1 - w1 * bn1;
    P[1] = n1 * PO[1] + t1 * PO[2] + m1 * PO[3];
    P[2] = n2 * PO[1] + t2 * PO[2] + m2 * PO[3];
    P[3] = n3 * PO[1] + t3 * PO[2] + m3 * PO[3];
    P[4] = SPI4 * (n1 * PO[5] + t1 * PO[6] + m1 * PO[7]);
    P[5] = SPI4 * (n2 * PO[5] + t2 * PO[6] + m2 * PO[7]);
    P[6] = SPI4 * (n3 * PO[5] + t3 * PO[6] + m3 * PO[7]);
    P[0] = PO[0];
    P[7] = PO[4];
}

/* predictor (kernel.cu:1611-1720): physical fluxes at the four half-slope
 * face states Q -/+ DX/2, Q +/- DY/2 (rho, p falling back to the cell value
 * when not positive), a step of st with the axisymmetric source terms. */
void o_gkr_predictor(const o_grid* g, const double* Q, const double* DX, const double* DY, double* QQ, double st,
                     double y)
{
    const double gam = g->gamma, dx = g->dx, dy = g->dy;
    double P[8];
    double PSx = 0, PSy = 0, PUx = 0, PUy = 0, PBx = 0, PBy = 0, PBz = 0;
    /* west (-DX/2, n = -x), east (+DX/2, +x), north (+DY/2, +y), south (-DY/2, -y) */
    const double* D[4] = {DX, DX, DY, DY};
    const double sg[4] = {-1.0, 1.0, 1.0, -1.0};
    const double nn[4][2] = {{-1.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}, {0.0, -1.0}};
    const double len[4] = {dy, dy, dx, dx};
    for (int f = 0; f < 4; ++f) {
        const double* d = D[f];
        double ro_ = sg[f] < 0 ? Q[0] - d[0] / 2.0 : Q[0] + d[0] / 2.0;
        double p_ = sg[f] < 0 ? Q[1] - d[1] / 2.0 : Q[1] + d[1] / 2.0;
        if (ro_ <= 0.0) ro_ = Q[0];
        if (p_ <= 0.0) p_ = Q[1];
        double q[7];
        for (int k = 2; k < 7; ++k) q[k] = sg[f] < 0 ? Q[k] - d[k] / 2.0 : Q[k] + d[k] / 2.0;
        o_potok(gam, ro_, p_, q[2], q[3], 0.0, q[4], q[5], q[6], P, nn[f][0], nn[f][1], 0.0);
        PSx = PSx + P[0] * len[f];
        PSy = PSy + P[7] * len[f];
        PUx = PUx + P[1] * len[f];
        PUy = PUy + P[2] * len[f];
        PBx = PBx + P[4] * len[f];
        PBy = PBy + P[5] * len[f];
        PBz = PBz + P[6] * len[f];
    }
    const double dV = dx * dy;
    /* kernel.cu:1698-1719 */
    QQ[0] = Q[0] - st * PSx / dV - st * Q[0] * Q[3] / y;
    if (QQ[0] <= 0) {
        diag_inc(&g_diag.clamp_rho);
        QQ[0] = Q[0];
    }
    QQ[2] = (Q[0] * Q[2] - st * (PUx) / dV - st * (Q[0] * Q[2] * Q[3] - Q[4] * Q[5] / CPI4) / y) / QQ[0];
    QQ[3] = (Q[0] * Q[3] - st * (PUy) / dV - st * (Q[0] * Q[3] * Q[3] + (GKR_KV(Q[6]) - GKR_KV(Q[5])) / CPI4) / y) /
            QQ[0];
    QQ[4] = (Q[4] - st * (PBx) / dV - st * (Q[3] * Q[4] - Q[5] * Q[2]) / y);
    QQ[5] = (Q[5] - st * (PBy) / dV);
    QQ[6] = (Q[6] - st * (PBz) / dV);
    const double u8 = gkr_u8(gam, Q[0], Q[1], Q[2], Q[3], 0.0, Q[4], Q[5], Q[6]);
    QQ[1] = (u8 - st * (PSy) / dV -
             st * (((u8 + Q[1] + GKR_KVV(Q[4], Q[5], Q[6]) / CPI8) * Q[3] -
                    Q[5] * GKR_SKK(Q[2], Q[3], 0.0, Q[4], Q[5], Q[6]) / CPI4) / y) -
             0.5 * QQ[0] * GKR_KVV(QQ[2], QQ[3], 0.0) - GKR_KVV(QQ[4], QQ[5], QQ[6]) / CPI8) * (gam - 1.0);
    if (QQ[1] <= 0) {
        diag_inc(&g_diag.clamp_p);
        QQ[1] = 0.000001;
    }
}

static inline void gkr_load(const o_d2* s, const o_d2* u, const o_d3* b, long k, double* q)
{
    q[0] = s[k].x; q[1] = s[k].y; q[2] = u[k].x; q[3] = u[k].y; q[4] = b[k].x; q[5] = b[k].y; q[6] = b[k].z;
}

static inline void gkr_copy(double* d, const double* a) { for (int k = 0; k < 7; ++k) d[k] = a[k]; }

/* mirror in x (vx, Bx, By negated), in y (vy, Bz negated), the outflow
 * floor (vx raised to 0.01 where negative), the inflow state */
static inline void gkr_mx(double* d, const double* a) { gkr_copy(d, a); d[2] = -a[2]; d[4] = -a[4]; d[5] = -a[5]; }
static inline void gkr_my(double* d, const double* a) { gkr_copy(d, a); d[3] = -a[3]; d[6] = -a[6]; }
static inline void gkr_out(double* d, const double* a) { gkr_copy(d, a); if (d[2] < 0) d[2] = 0.01; }
static inline void gkr_in(double* d, double gam, const double* bsrc)
{
    d[0] = 1.0; d[1] = 1.0 / gam; d[2] = 0.0; d[3] = 0.0; d[4] = bsrc[4]; d[5] = bsrc[5]; d[6] = bsrc[6];
}

/* the 13-point stencil of kernel.cu:2678-2957, st[k] for the reference's
 * s_k / u_k / b_k (k = 1..13) */
void o_gkr_stencil(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, int n, int m, double st[14][7])
{
    const int N = g->nx, M = g->ny;
    const double gam = g->gamma;
#define AT(nn, mm, q) gkr_load(s, u, b, (long)(mm) * N + (nn), q)
    AT(n, m, st[1]);
    if (m == M - 1) gkr_in(st[5], gam, st[1]); else AT(n, m + 1, st[5]);
    if (n == N - 1) gkr_out(st[2], st[1]); else AT(n + 1, m, st[2]);
    if (n == 0) gkr_mx(st[4], st[1]); else AT(n - 1, m, st[4]);
    if (m == 0) gkr_my(st[3], st[1]); else AT(n, m - 1, st[3]);
    /* kernel.cu:2766-2781 */
    if (n == N - 1) gkr_copy(st[10], st[5]);
    else if (m == M - 1) gkr_in(st[10], gam, st[2]);
    else AT(n + 1, m + 1, st[10]);
    /* kernel.cu:2784-2840 */
    if (m == 0) {
        gkr_my(st[12], st[4]);
        gkr_my(st[11], st[2]);
    } else {
        if (n == 0) gkr_mx(st[12], st[3]); else AT(n - 1, m - 1, st[12]);
        if (n == N - 1) gkr_out(st[11], st[3]); else AT(n + 1, m - 1, st[11]);
    }
    /* kernel.cu:2842-2870 */
    if (n == 0) gkr_mx(st[13], st[5]);
    else if (m == M - 1) gkr_in(st[13], gam, st[4]);
    else AT(n - 1, m + 1, st[13]);
    /* kernel.cu:2872-2957 */
    if (n >= N - 2) gkr_out(st[6], st[2]); else AT(n + 2, m, st[6]);
    if (m >= M - 2) gkr_copy(st[9], st[5]); else AT(n, m + 2, st[9]);
    if (n == 0 || n == 1) gkr_mx(st[8], st[2]); else AT(n - 2, m, st[8]);
    if (m == 1) gkr_my(st[7], st[5]);
    else if (m == 0) gkr_my(st[7], st[3]);
    else AT(n, m - 2, st[7]);
#undef AT
}

/* dx * minmod((a - c)/dx, (c - b)/dx) per component, the vy slot taking the
 * pressure slope (kernel.cu:2962-3038) */
static inline void gkr_slope(const double* hi, const double* mid, const double* lo, double h, double* D)
{
    static const int src[7] = {0, 1, 2, 1, 4, 5, 6};
    for (int k = 0; k < 7; ++k) {
        const int c = src[k];
        D[k] = h * o_minmod((hi[c] - mid[c]) / h, (mid[c] - lo[c]) / h);
    }
}

void o_gkr_slopes(const o_grid* g, double st[14][7], double DX[6][7], double DY[6][7])
{
    const double dx = g->dx, dy = g->dy;
    gkr_slope(st[2], st[1], st[4], dx, DX[1]);
    gkr_slope(st[6], st[2], st[1], dx, DX[2]);
    gkr_slope(st[1], st[4], st[8], dx, DX[4]);
    gkr_slope(st[10], st[5], st[13], dx, DX[5]);
    gkr_slope(st[11], st[3], st[12], dx, DX[3]);
    gkr_slope(st[5], st[1], st[3], dy, DY[1]);
    gkr_slope(st[9], st[5], st[1], dy, DY[5]);
    gkr_slope(st[13], st[4], st[12], dy, DY[4]);
    gkr_slope(st[10], st[2], st[11], dy, DY[2]);
    gkr_slope(st[1], st[3], st[7], dy, DY[3]);
}

int o_kernel_gkr(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, o_d2* s2, o_d2* u2, o_d3* b2,
                 double* T, const double* T_do, int method)
{
    if (method != 2 && method != 3) return -1;
    if (!(g->y0 > 0.0)) return -1;
    const double dx = g->dx, dy = g->dy, gam = g->gamma;
    const double Tdo = *T_do;
    double tglob = *T;
#pragma omp parallel for schedule(static) reduction(min : tglob)
    for (int m = 0; m < g->ny; ++m) {
        for (int n = 0; n < g->nx; ++n) {
            const long k = (long)m * g->nx + n;
            const double x = g->x0 + n * dx, y = g->y0 + m * dy;
            if (sqrt(x * x + y * y) < g->r_frozen) {
                s2[k] = s[k]; u2[k] = u[k]; b2[k] = b[k];
                continue;
            }
            double st[14][7], DX[6][7], DY[6][7], QQ[6][7];
            o_gkr_stencil(g, s, u, b, n, m, st);
            o_gkr_slopes(g, st, DX, DY);
            /* kernel.cu:3089-3093: cells 1..5 at y, y, y - dy, y, y + dy */
            const double yk[6] = {0.0, y, y, y - dy, y, y + dy};
            for (int c = 1; c <= 5; ++c) o_gkr_predictor(g, st[c], DX[c], DY[c], QQ[c], Tdo / 2.0, yk[c]);
            double tmin = 1000;
            double PSx = 0, PSy = 0, PUx = 0, PUy = 0, PBx = 0, PBy = 0, PBz = 0;
            /* kernel.cu:3105-3238: east, south, west, north faces */
            const int nb[4] = {2, 3, 4, 5};
            const double* Din[4] = {DX[1], DY[1], DX[1], DY[1]};
            const double* Dnb[4] = {DX[2], DY[3], DX[4], DY[5]};
            const double sgn_in[4] = {1.0, -1.0, -1.0, 1.0};
            const double nn[4][2] = {{1.0, 0.0}, {0.0, -1.0}, {-1.0, 0.0}, {0.0, 1.0}};
            const double rad[4] = {dx, dy, dx, dy};
            const double len[4] = {dy, dx, dy, dx};
            for (int f = 0; f < 4; ++f) {
                const int c = nb[f];
                double L[7], R[7];
                for (int q = 0; q < 7; ++q) {
                    const double m1 = (st[1][q] + QQ[1][q]) / 2.0, m2 = (st[c][q] + QQ[c][q]) / 2.0;
                    L[q] = sgn_in[f] > 0 ? m1 + Din[f][q] / 2.0 : m1 - Din[f][q] / 2.0;
                    R[q] = sgn_in[f] > 0 ? m2 - Dnb[f][q] / 2.0 : m2 + Dnb[f][q] / 2.0;
                    if (q < 2) {
                        if (L[q] <= 0.0) L[q] = m1;
                        if (R[q] <= 0.0) R[q] = m2;
                    }
                }
                double P[8];
                const double t = o_hlld(gam, g->krit, L[0], L[1], L[2], L[3], 0.0, L[4], L[5], L[6], R[0], R[1],
                                        R[2], R[3], 0.0, R[4], R[5], R[6], P, nn[f][0], nn[f][1], 0.0, rad[f],
                                        method);
                tmin = tmin <= t ? tmin : t; /* my_min, kernel.cu:102 */
                PSx = PSx + P[0] * len[f];
                PSy = PSy + P[7] * len[f];
                PUx = PUx + P[1] * len[f];
                PUy = PUy + P[2] * len[f];
                PBx = PBx + P[4] * len[f];
                PBy = PBy + P[5] * len[f];
                PBz = PBz + P[6] * len[f];
            }
            if (tmin < tglob) tglob = tmin;
            /* kernel.cu:3247-3283: the corrector from the mean of the cell and
             * its predictor, dt/2, Powell's term zero */
            const double dV = dx * dy, h = Tdo / 2.0;
            double Qb[7];
            for (int q = 0; q < 7; ++q) Qb[q] = (st[1][q] + QQ[1][q]) / 2.0;
            double r2 = Qb[0] - h * PSx / dV - h * Qb[0] * Qb[3] / y;
            if (r2 <= 0) {
                diag_inc(&g_diag.clamp_rho);
                r2 = Qb[0];
            }
            const double vx2 = (Qb[0] * Qb[2] - h * (PUx) / dV - h * (Qb[0] * Qb[2] * Qb[3] - Qb[4] * Qb[5] / CPI4) / y) / r2;
            const double vy2 = (Qb[0] * Qb[3] - h * (PUy) / dV -
                                h * (Qb[0] * Qb[3] * Qb[3] + (GKR_KV(Qb[6]) - GKR_KV(Qb[5])) / CPI4) / y) / r2;
            const double bx2 = (Qb[4] - h * (PBx) / dV - h * (Qb[3] * Qb[4] - Qb[5] * Qb[2]) / y);
            const double by2 = (Qb[5] - h * (PBy) / dV);
            const double bz2 = (Qb[6] - h * (PBz) / dV);
            const double u8 = gkr_u8(gam, Qb[0], Qb[1], Qb[2], Qb[3], 0.0, Qb[4], Qb[5], Qb[6]);
            double p2 = (u8 - h * (PSy) / dV -
                         h * (((u8 + Qb[1] + GKR_KVV(Qb[4], Qb[5], Qb[6]) / CPI8) * Qb[3] -
                               Qb[5] * GKR_SKK(Qb[2], Qb[3], 0.0, Qb[4], Qb[5], Qb[6]) / CPI4) / y) -
                         0.5 * r2 * GKR_KVV(vx2, vy2, 0.0) - GKR_KVV(bx2, by2, bz2) / CPI8) * (gam - 1.0);
            if (p2 <= 0) {
                diag_inc(&g_diag.clamp_p);
                p2 = 0.000001;
            }
            s2[k].x = r2; s2[k].y = p2;
            u2[k].x = vx2; u2[k].y = vy2;
            b2[k].x = bx2; b2[k].y = by2; b2[k].z = bz2;
        }
    }
    *T = tglob;
    return 0;
}

/* kernel.cu:1722-1733 (the every-5000-steps print is left out) */
void o_funk_time(double* T, double* T_do, double* TT, int* i)
{
    *T_do = *T;
    *TT = *TT + *T_do;
    *T = 10000000;
    *i = *i + 1;
}

/* ------------------------------------------------------------------------ */
/* native path (SoA conserved)                                                */
/* ------------------------------------------------------------------------ */

static inline double cell_wavespeed(const o_grid* g, const prim* q)
{
    const double idx = 1.0 / g->dx, idy = 1.0 / g->dy;
    /* sound speed (Euler) or the fastest magnetosonic speed bound
     * sqrt((gamma p + |B_HL|^2) / rho) with B_HL = B_gauss / sqrt(4 pi) */
    double b2 = (q->bx * q->bx + q->by * q->by + q->bz * q->bz) / CPI4;
    double c = sqrt((g->gamma * q->p + b2) / q->r);
    return (fabs(q->u) + c) * idx + (fabs(q->v) + c) * idy;
}

static int fields_of(int physics) { return physics == O_PHYS_MHD ? 7 : 4; }

int o_stage(const o_grid* g, const o_scheme* sc, const double* Uin, const double* Ubase,
            double* Uout, double a, double b, double dt, double* wmax)
{
    const int mhd = sc->physics == O_PHYS_MHD;
    if (!mhd && sc->flux != O_FLUX_HLLC && sc->flux != O_FLUX_EXACT && sc->flux != O_FLUX_HLL &&
        sc->flux != O_FLUX_HLLC_KOROLKOV && sc->flux != O_FLUX_HLLC_2D)
        return -1;
    if (mhd && sc->flux != O_FLUX_HLLC_MHD && sc->flux != O_FLUX_HLLD) return -1;
    for (int s = 0; s < 4; ++s)
        if (g->bc[s] == O_BC_FROZEN) return -1;
    const int plain = (a == 0.0 && b == 1.0);
    src_t S = {g, 1, mhd, NULL, NULL, NULL, Uin, 0};
    const double dx = g->dx, dy = g->dy, gam = g->gamma, dV = dx * dy;
    const long plane = (long)g->nx * g->ny;
    const int nf = fields_of(sc->physics);
    const int method = mhd ? (sc->flux == O_FLUX_HLLD ? 3 : 2)
                           : (sc->flux == O_FLUX_EXACT ? 2 : (sc->flux == O_FLUX_HLL ? 0 :
                              (sc->flux == O_FLUX_HLLC_KOROLKOV ? 3 : (sc->flux == O_FLUX_HLLC_2D ? 4 : 1))));
    double wm = 0.0;
#pragma omp parallel for schedule(static) reduction(max : wm)
    for (int j = 0; j < g->ny; ++j) {
        for (int i = 0; i < g->nx; ++i) {
            long k = (long)j * g->nx + i;
            stencil st;
            faces F;
            gather(&S, i, j, sc->order, &st);
            reconstruct(g, i, j, sc->order, mhd, &st, &F);
            double un[7];
            if (!mhd) {
                euler_sum E = euler_cell_fluxes(g, method, &F, 0);
                /* forward-Euler update in conserved form, kernel.cu:1859-1870 */
                un[0] = Uin[k] - (dt / dV) * E.PS.x;
                un[1] = Uin[plane + k] - (dt / dV) * E.PU.x;
                un[2] = Uin[2 * plane + k] - (dt / dV) * E.PU.y;
                un[3] = Uin[3 * plane + k] - (dt / dV) * E.PS.y;
            } else {
                mhd_sum m = mhd_cell_fluxes(g, method, &F, &st);
                const prim* c = &st.c;
                /* kernel.cu:2619-2648 (Cartesian), conserved form */
                un[0] = Uin[k] - dt * m.PS.x / dV;
                un[1] = Uin[plane + k] - dt * (m.PU.x + (c->bx / CPI4) * m.Pdiv) / dV;
                un[2] = Uin[2 * plane + k] - dt * (m.PU.y + (c->by / CPI4) * m.Pdiv) / dV;
                un[3] = Uin[3 * plane + k] - dt * (m.PB.x + c->u * m.Pdiv) / dV;
                un[4] = Uin[4 * plane + k] - dt * (m.PB.y + c->v * m.Pdiv) / dV;
                un[5] = Uin[5 * plane + k] - dt * (m.PB.z) / dV;
                double skk = c->u * c->bx + c->v * c->by;
                un[6] = Uin[6 * plane + k] - dt * (m.PS.y + (skk / CPI4) * m.Pdiv) / dV;
            }
            double out[7];
            for (int f = 0; f < nf; ++f)
                out[f] = plain ? un[f] : a * Ubase[f * plane + k] + b * un[f];
            /* clamps: rho (kernel.cu:1862-1866 Euler -> 1e-4; 2144-2148 MHD -> old rho),
             * p (kernel.cu:1871-1874 -> 1e-6) */
            if (out[0] <= 0) {
                diag_inc(&g_diag.clamp_rho);
                out[0] = mhd ? Uin[k] : 0.0001;
            }
            {
                double inv = 1.0 / out[0];
                double ux = out[1] * inv, uy = out[2] * inv;
                double ke = 0.5 * (out[1] * ux + out[2] * uy);
                double me = mhd ? (out[3] * out[3] + out[4] * out[4] + out[5] * out[5]) / CPI8 : 0.0;
                int ie = nf - 1;
                double p = (gam - 1.0) * (out[ie] - ke - me);
                if (p <= 0) {
                    diag_inc(&g_diag.clamp_p);
                    out[ie] = 0.000001 / (gam - 1.0) + ke + me;
                }
            }
            for (int f = 0; f < nf; ++f) Uout[f * plane + k] = out[f];
            if (wmax) {
                src_t So = {g, 1, mhd, NULL, NULL, NULL, Uout, 0};
                prim q = cons_cell_to_prim(&So, k);
                double w = cell_wavespeed(g, &q);
                if (w > wm) wm = w;
            }
        }
    }
    if (wmax) *wmax = wm;
    return 0;
}

double o_wavespeed_max(const o_grid* g, const o_scheme* sc, const double* U)
{
    const int mhd = sc->physics == O_PHYS_MHD;
    src_t S = {g, 1, mhd, NULL, NULL, NULL, U, 0};
    const long n = (long)g->nx * g->ny;
    double wm = 0.0;
#pragma omp parallel for schedule(static) reduction(max : wm)
    for (long k = 0; k < n; ++k) {
        prim q = cons_cell_to_prim(&S, k);
        double w = cell_wavespeed(g, &q);
        if (w > wm) wm = w;
    }
    return wm;
}

long o_run(const o_grid* g, const o_scheme* sc, int integrator, double* U, double* tmp,
           long nsteps, double t_end, double* t)
{
    const long plane = (long)g->nx * g->ny;
    const int nf = fields_of(sc->physics);
    double* U1 = tmp;
    double* U2 = tmp + (long)nf * plane;
    double wmax = o_wavespeed_max(g, sc, U);
    long steps = 0;
    double tt = *t;
    while (nsteps <= 0 || steps < nsteps) {
        if (t_end > 0 && !(tt < t_end)) break;
        double dt = sc->cfl / wmax;
        if (t_end > 0 && dt > t_end - tt) dt = t_end - tt;
        if (integrator == 1) {
            if (o_stage(g, sc, U, NULL, U1, 0.0, 1.0, dt, &wmax)) return -1;
            memcpy(U, U1, sizeof(double) * nf * plane);
        } else {
            if (o_stage(g, sc, U, NULL, U1, 0.0, 1.0, dt, NULL)) return -1;
            if (o_stage(g, sc, U1, U, U2, 0.5, 0.5, dt, &wmax)) return -1;
            memcpy(U, U2, sizeof(double) * nf * plane);
        }
        tt = tt + dt;
        ++steps;
    }
    *t = tt;
    return steps;
}

void o_prim_to_cons(const o_grid* g, int physics, const o_d2* s, const o_d2* u, const o_d3* b, double* U)
{
    const long plane = (long)g->nx * g->ny;
    const double gam = g->gamma;
    for (long k = 0; k < plane; ++k) {
        double r = s[k].x, p = s[k].y, vx = u[k].x, vy = u[k].y;
        U[k] = r;
        U[plane + k] = r * vx;
        U[2 * plane + k] = r * vy;
        if (physics == O_PHYS_MHD) {
            U[3 * plane + k] = b[k].x;
            U[4 * plane + k] = b[k].y;
            U[5 * plane + k] = b[k].z;
            U[6 * plane + k] = p / (gam - 1.0) + 0.5 * r * (vx * vx + vy * vy) +
                               (b[k].x * b[k].x + b[k].y * b[k].y + b[k].z * b[k].z) / CPI8;
        } else {
            U[3 * plane + k] = p / (gam - 1.0) + 0.5 * r * (vx * vx + vy * vy);
        }
    }
}

void o_cons_to_prim(const o_grid* g, int physics, const double* U, o_d2* s, o_d2* u, o_d3* b)
{
    src_t S = {g, 1, physics == O_PHYS_MHD, NULL, NULL, NULL, U, 0};
    const long plane = (long)g->nx * g->ny;
    for (long k = 0; k < plane; ++k) {
        prim q = cons_cell_to_prim(&S, k);
        s[k].x = q.r; s[k].y = q.p; u[k].x = q.u; u[k].y = q.v;
        if (b) { b[k].x = q.bx; b[k].y = q.by; b[k].z = q.bz; }
    }
}

/* ------------------------------------------------------------------ on-disk formats
 * One cell at a time with stdio, as the reference's host driver writes with
 * std::ofstream (whose default float output is printf "%g"). */

/* instable_for_inner_b3.txt writer (kernel.cu:4252-4260) */
int o_write_state_text(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, const char* path)
{
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    const long K = (long)g->nx * g->ny;
    for (long k = 0; k < K; ++k) {
        int n = (int)(k % g->nx), m = (int)(k / g->nx);
        double y = g->y0 + m * g->dy;
        double x = g->x0 + n * g->dx;
        fprintf(f, "%g %g %g %g %g %g %g %g %g\n", x, y, s[k].x, s[k].y, u[k].x, u[k].y,
                b ? b[k].x : 0.0, b ? b[k].y : 0.0, b ? b[k].z : 0.0);
    }
    return fclose(f) == 0 ? 0 : -1;
}

/* restart reader (kernel.cu:3944-3969): c1 c2 a1 ... a7 per cell */
int o_read_state_text(const o_grid* g, const char* path, o_d2* s, o_d2* u, o_d3* b)
{
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    const long K = (long)g->nx * g->ny;
    for (long k = 0; k < K; ++k) {
        double c1, c2, a[7];
        if (fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf", &c1, &c2, &a[0], &a[1], &a[2], &a[3], &a[4], &a[5],
                   &a[6]) != 9) {
            fclose(f);
            return -1;
        }
        s[k].x = a[0]; s[k].y = a[1];
        u[k].x = a[2]; u[k].y = a[3];
        if (b) { b[k].x = a[4]; b[k].y = a[5]; b[k].z = a[6]; }
    }
    fclose(f);
    return 0;
}

/* save_b3.dat (kernel.cu:4261-4267) */
int o_write_state_b3(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, const char* path)
{
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    const long K = (long)g->nx * g->ny;
    int ok = 1;
    for (long k = 0; k < K && ok; ++k) {
        int n = (int)(k % g->nx), m = (int)(k / g->nx);
        double y = g->y0 + m * g->dy;
        double x = g->x0 + n * g->dx;
        ok = fwrite(&x, 8, 1, f) == 1 && fwrite(&y, 8, 1, f) == 1 && fwrite(&s[k].x, 8, 1, f) == 1 &&
             fwrite(&s[k].y, 8, 1, f) == 1 && fwrite(&u[k].x, 8, 1, f) == 1 && fwrite(&u[k].y, 8, 1, f) == 1 &&
             fwrite(&b[k].z, 8, 1, f) == 1;
    }
    return (fclose(f) == 0 && ok) ? 0 : -1;
}

/* print_file_mini2 (kernel.cu:3724-3808) with b, print_file_mini
 * (kernel.cu:3682-3722) without; x_max / y_max are the domain's edges */
int o_write_tecplot(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, int nmin,
                    const double* solution_time, const char* path)
{
    if (nmin < 1) return -1;
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    const int N = g->nx, M = g->ny;
    const long K = (long)N * M;
    const int nn = (N + nmin - 1) / nmin, mm = (M + nmin - 1) / nmin;
    const double x_min = g->x0, y_min = g->y0;
    const double x_max = g->x0 + (N - 0.5) * g->dx, y_max = g->y0 + (M - 0.5) * g->dy;
    if (b)
        fprintf(f, "TITLE = \"HP\"  VARIABLES = \"X\", \"Y\", \"Ro\", \"P\", \"Vx\", \"Vy\", \"Bx\", \"By\", \"Bz\", "
                   "\"Max\", \"T\", ZONE T = \"HP\", N = %d , E = %d, F = FEPOINT, ET = quadrilateral",
                nn * mm, (nn - 1) * (mm - 1));
    else
        fprintf(f, "TITLE = \"HP\"  VARIABLES = \"X\", \"Y\", \"Ro\", \"P\", \"Vx\", \"Vy\", \"Max\", \"T\", "
                   "ZONE T = \"HP\", N = %d , E = %d, F = FEPOINT, ET = quadrilateral",
                nn * mm, (nn - 1) * (mm - 1));
    if (solution_time) fprintf(f, ", SOLUTIONTIME = %g", *solution_time);
    fprintf(f, "\n");
    for (long k = 0; k < K; k++) {
        int n = (int)(k % N);
        int m = (int)((k - n) / N);
        if ((n % nmin != 0) || (m % nmin != 0)) continue;
        double y = M > 1 ? y_min + m * (y_max - y_min) / (M - 1) : y_min;
        double x = N > 1 ? x_min + n * (x_max - x_min) / (N - 1) : x_min;
        double Max = 0.0, Temp = 0.0;
        if (s[k].x > 0.0) {
            Max = sqrt((u[k].x * u[k].x + u[k].y * u[k].y) / (g->gamma * s[k].y / s[k].x));
            Temp = s[k].y / s[k].x;
        }
        if (b)
            fprintf(f, "%g %g %g %g %g %g %g %g %g %g %g\n", x, y, s[k].x, s[k].y, u[k].x, u[k].y, b[k].x, b[k].y,
                    b[k].z, Max, Temp);
        else
            fprintf(f, "%g %g %g %g %g %g %g %g\n", x, y, s[k].x, s[k].y, u[k].x, u[k].y, Max, Temp);
    }
    for (int k = 0; k < nn * mm; k = k + 1) {
        int n = k % nn;
        int m = (k - n) / nn;
        if ((m < mm - 1) && (n < nn - 1))
            fprintf(f, "%d %d %d %d\n", m * nn + n + 1, m * nn + n + 2, (m + 1) * nn + n + 2, (m + 1) * nn + n + 1);
    }
    return fclose(f) == 0 ? 0 : -1;
}

/* End-of-run derived outputs (kernel.cu:4231-4460), one cell at a time as
 * the reference's loops write them. pi as Header.h:18. */
#define O_PI 3.14159265358979323846

/* param_for_texplot_mini.txt (kernel.cu:4384-4460) */
int o_write_tecplot_derived(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, int nmin,
                            const char* path)
{
    if (nmin < 1) return -1;
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    const int N = g->nx, M = g->ny;
    const long K = (long)N * M;
    const int nn = (N + nmin - 1) / nmin, mm = (M + nmin - 1) / nmin;
    fprintf(f, "TITLE = \"HP\"  VARIABLES = \"X\", \"Y\", \"Ro\", \"P\", \"Vx\", \"Vy\", \"Bx\", \"By\", \"Bz\", "
               "\"Max\", \"T\", \"PP\", \"Fx\", \"Fy\", \"zav\",  ZONE T = \"HP\", N = %d , E = %d, F = FEPOINT, "
               "ET = quadrilateral\n", nn * mm, (nn - 1) * (mm - 1));
    for (long k = 0; k < K; k++) {
        int n = (int)(k % N);
        int m = (int)((k - n) / N);
        double y = g->y0 + m * g->dy;
        double x = g->x0 + n * g->dx;
        if ((n % nmin != 0) || (m % nmin != 0)) continue;
        long kk = (long)(m + 1) * N + n, kk2 = (long)(m - 1) * N + n;
        long e = (long)m * N + n + 1, w = (long)m * N + n - 1;
        double zav = 0.0, Fx = 0.0, Fy = 0.0;
        if (m < M - 1 && n < N - 1) {
            double dAr, dAz;
            if (n > 0) dAz = (b[e].z - b[w].z) / (2.0 * g->dx);
            else dAz = (b[e].z - b[k].z) / (g->dx);
            if (m > 0) dAr = (b[kk].z - b[kk2].z) / (2.0 * g->dy);
            else dAr = (b[kk].z - b[k].z) / (g->dy);
            Fx = b[k].z * dAz / (4.0 * O_PI);
            Fy = (-b[k].z * b[k].z / y - b[k].z * dAr) / (4.0 * O_PI);
            if (n > 0 && m > 0)
                zav = (u[e].y - u[w].y) / (2.0 * g->dx) - (u[kk].x - u[kk2].x) / (2.0 * g->dy);
        }
        double Max = 0.0, Temp = 0.0;
        if (s[k].x > 0.0) {
            Max = sqrt((u[k].x * u[k].x + u[k].y * u[k].y) / (g->gamma * s[k].y / s[k].x));
            Temp = s[k].y / s[k].x;
        }
        fprintf(f, "%g %g %g %g %g %g %g %g %g %g %g %g %g %g %g\n", x, y, s[k].x, s[k].y, u[k].x, u[k].y, b[k].x,
                b[k].y, b[k].z, Max, Temp, s[k].y + (b[k].z * b[k].z) / (8.0 * O_PI), Fx, Fy, zav);
    }
    for (int k = 0; k < nn * mm; k = k + 1) {
        int n = k % nn;
        int m = (k - n) / nn;
        if ((m < mm - 1) && (n < nn - 1))
            fprintf(f, "%d %d %d %d\n", m * nn + n + 1, m * nn + n + 2, (m + 1) * nn + n + 2, (m + 1) * nn + n + 1);
    }
    return fclose(f) == 0 ? 0 : -1;
}

/* param_y=0.txt (kernel.cu:4283-4327) and param_x=0.txt (kernel.cu:4330-4374) */
int o_write_line_cuts(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, const char* path_y0,
                      const char* path_x0)
{
    const int N = g->nx, M = g->ny;
    if (path_y0) {
        FILE* f = fopen(path_y0, "w");
        if (!f) return -1;
        fprintf(f, "TITLE = \"HP\"  VARIABLES = \"X\", \"Ro\", \"P\", \"Vx\", \"Bz\", \"Max\", \"T\", \"PP\", \"Fx\", "
                   "\"Fy\", \"dP\", ZONE T = \"HP\"\n");
        for (int k = 5; k < N - 1; k++) {
            int n = k % N;
            int m = (k - n) / N;
            double y = g->y0 + m * g->dy;
            double x = g->x0 + n * g->dx;
            long kk = (long)(m + 1) * N + n;
            double Max = 0.0, Temp = 0.0;
            if (s[k].x > 0.0) {
                Max = sqrt((u[k].x * u[k].x + u[k].y * u[k].y) / (g->gamma * s[k].y / s[k].x));
                Temp = s[k].y / s[k].x;
            }
            long e = (long)m * N + n + 1, w = (long)m * N + n - 1;
            double Fx = 0.0, Fy = 0.0;
            if (m < M - 1 && n < N - 1) {
                double dAz = n > 0 ? (b[e].z - b[w].z) / (2.0 * g->dx) : (b[e].z - b[k].z) / (g->dx);
                double dAr = (b[kk].z - b[k].z) / (g->dy);
                Fx = b[k].z * dAz / (4.0 * O_PI);
                Fy = (-b[k].z * b[k].z / y - b[k].z * dAr) / (4.0 * O_PI);
            }
            fprintf(f, "%g %g %g %g %g %g %g %g %g %g %g\n", x, s[k].x, s[k].y, u[k].x, b[k].z, Max, Temp,
                    s[k].y + (b[k].z * b[k].z) / (8.0 * O_PI), Fx, Fy,
                    ((s[e].y + (b[e].z * b[e].z) / (8.0 * O_PI)) - (s[k].y + (b[k].z * b[k].z) / (8.0 * O_PI))) / g->dx);
        }
        if (fclose(f) != 0) return -1;
    }
    if (path_x0) {
        FILE* f = fopen(path_x0, "w");
        if (!f) return -1;
        fprintf(f, "TITLE = \"HP\"  VARIABLES = \"Y\", \"Ro\", \"P\", \"Vy\", \"Bz\", \"Max\", \"T\", \"PP\", \"Fx\", "
                   "\"Fy\", \"dP\", ZONE T = \"HP\"\n");
        for (int m = 5; m < M - 1; m++) {
            int n = 0;
            double y = g->y0 + m * g->dy;
            long kk = (long)(m + 1) * N;
            long e = (long)m * N + 0 + 1;
            long k = (long)m * N + 0;
            double Max = 0.0, Temp = 0.0;
            if (s[k].x > 0.0) {
                Max = sqrt((u[k].x * u[k].x + u[k].y * u[k].y) / (g->gamma * s[k].y / s[k].x));
                Temp = s[k].y / s[k].x;
            }
            double Fx = 0.0, Fy = 0.0;
            if (m < M - 1 && n < N - 1) {
                double dAz = (b[e].z - b[k].z) / (g->dx);
                double dAr = (b[kk].z - b[k].z) / (g->dy);
                Fx = b[k].z * dAz / (4.0 * O_PI);
                Fy = (-b[k].z * b[k].z / y - b[k].z * dAr) / (4.0 * O_PI);
            }
            fprintf(f, "%g %g %g %g %g %g %g %g %g %g %g\n", y, s[k].x, s[k].y, u[k].y, b[k].z, Max, Temp,
                    s[k].y + (b[k].z * b[k].z) / (8.0 * O_PI), Fx, Fy,
                    ((s[kk].y + (b[kk].z * b[kk].z) / (8.0 * O_PI)) - (s[k].y + (b[k].z * b[k].z) / (8.0 * O_PI))) / g->dy);
        }
        if (fclose(f) != 0) return -1;
    }
    return 0;
}
