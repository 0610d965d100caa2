/*
 * gd2_oracle.h -- CPU restatement of the reference's explicit finite-volume
 * step (SergeyMSU/Golikov_2d). TEST INFRASTRUCTURE ONLY: the parity tests,
 * the smoke check and the CPU-baseline leg of bench.py link this library as
 * the checker. Nothing in the product path may call it.
 *
 * Every function cites the reference lines it restates (file:line in the
 * reference repository, files read as cp1251).
 */
#ifndef GD2_ORACLE_H
#define GD2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y; } o_d2;
typedef struct { double x, y, z; } o_d3;

/* Side boundary kinds (same numbering as the product's gd2_bc). */
enum { O_BC_TRANSMISSIVE = 0, O_BC_PERIODIC = 1, O_BC_REFLECTIVE = 2, O_BC_FROZEN = 3,
       /* the production ghost rules of add2 / Kernel_TVD (one side each) */
       O_BC_SYMMETRY = 4, O_BC_AXIS = 5, O_BC_OUTFLOW = 6, O_BC_INFLOW = 7,
       /* Kernel_TVD_old's west ghost (kernel.cu:3323-3337, 3365-3378) */
       O_BC_WEST_INFLOW = 8 };

/* Runtime grid/physics descriptor: the reference has these as macros
 * (Header.h:3-16 gamma/krit, kernel.cu:11-20 N, M, dx, dy, x_min, y_min). */
typedef struct {
    int nx, ny;          /* N (x, fastest), M (y) */
    double dx, dy;
    double x0, y0;       /* centre of cell (0,0): x_min, y_min */
    double gamma;        /* ga / ggg */
    double krit;         /* CFL factor of the reference-layout steps */
    int bc[4];           /* west, east, south, north */
    double r_frozen;     /* MHD reference steps: cells inside this radius keep their state */
    int axisym;          /* MHD reference steps: axisymmetric (y = r) source terms */
    double m_inf;        /* M_inf (Header.h:24) of the WEST_INFLOW side */
} o_grid;

/* Scheme selection for the native (SoA conserved) path. */
enum { O_PHYS_EULER = 0, O_PHYS_MHD = 1 };
enum { O_FLUX_HLLC = 1, O_FLUX_EXACT = 2, O_FLUX_HLLC_MHD = 3, O_FLUX_HLLD = 4, O_FLUX_HLL = 5,
       O_FLUX_HLLC_KOROLKOV = 6, O_FLUX_HLLC_2D = 7 };

typedef struct {
    int physics;         /* O_PHYS_* */
    int flux;            /* O_FLUX_* */
    int order;           /* 1 = piecewise constant, 2 = minmod MUSCL */
    double cfl;          /* Courant number of the native path */
} o_scheme;

/* Diagnostics counters (iteration caps hit, clamps applied). */
typedef struct {
    long newton_cap_hits, bisect_cap_hits, clamp_rho, clamp_p, bad_branch;
} o_diag;

void o_diag_reset(void);
void o_diag_get(o_diag* out);

/* ---- pointwise helpers (kernel.cu:51-100) ---- */
double o_minmod(double x, double y);
double o_linear(double x1, double t1, double x2, double t2, double x3, double t3, double y);
void   o_linear2(double x1, double t1, double x2, double t2, double x3, double t3,
                 double y1, double y2, double* A, double* B);

/* ---- interface fluxes; each returns the face time step krit*rad/S ---- */
double o_hll(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
             double n1, double n2, o_d2* Ps, o_d2* Pu, double rad);
double o_hllc(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
              double n1, double n2, o_d2* Ps, o_d2* Pu, double rad);
double o_godunov(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
                 double n1, double n2, o_d2* Ps, o_d2* Pu, double rad);
/* the remaining HLLC variants (kernel.cu:1167, 1349; Gas_Dinamic.cu:265, 359,
 * 465); the Q variants also return the passive-scalar flux in *PQ (may be NULL) */
double o_hllc_aleksashov2(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
                          double n1, double n2, o_d2* Ps, o_d2* Pu, double rad);
double o_hllc_korolkov(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
                       double n1, double n2, o_d2* Ps, o_d2* Pu, double rad);
double o_hllcq_korolkov_2d(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru, double Q_L,
                           double Q_R, double n1, double n2, o_d2* Ps, o_d2* Pu, double* PQ, double rad);
double o_hllcq_aleksashov(double gamma, double krit, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru, double Q_L,
                          double Q_R, double n1, double n2, o_d2* Ps, o_d2* Pu, double* PQ, double rad);
/* star pressure / velocity of the exact solver, in the caller's frame along n */
void   o_godunov_star(double gamma, o_d2 Ls, o_d2 Lu, o_d2 Rs, o_d2 Ru,
                      double n1, double n2, double* pstar, double* ustar, int* pattern);
/* MHD flux: P[8] = (F_rho, F_mx, F_my, F_mz, F_Bx, F_By, F_Bz, F_E), Gaussian B.
 * metod 2 = HLLC-MHD, 3 = HLLD; any other value returns -1 and zeroes P. */
double o_hlld(double gamma, double krit,
              double ro_L, double p_L, double v1_L, double v2_L, double v3_L,
              double Bx_L, double By_L, double Bz_L,
              double ro_R, double p_R, double v1_R, double v2_R, double v3_R,
              double Bx_R, double By_R, double Bz_R,
              double* P, double n1, double n2, double n3, double rad, int metod);

/* ---- reference-layout steps (AoS primitives, lagged dt folded into *T) ---- */
/* Euler, first order: Ker_Dekard (kernel.cu:1736-1875). method 1 = HLLC, 2 = exact. */
int o_ker_dekard(const o_grid* g, const o_d2* s, const o_d2* u, o_d2* s2, o_d2* u2,
                 double* T, const double* T_do, int method);
/* Euler, minmod MUSCL: Cartesian part of Kernel_TVD_old (kernel.cu:3286-3666). */
int o_kernel_tvd_old(const o_grid* g, const o_d2* s, const o_d2* u, o_d2* s2, o_d2* u2,
                     double* T, const double* T_do, int method);
/* MHD first order + Powell source: Cartesian part of add2 (kernel.cu:1877-2196). */
int o_add2(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b,
           o_d2* s2, o_d2* u2, o_d3* b2, double* T, const double* T_do, int i, int method);
/* MHD minmod MUSCL + Powell source: Cartesian part of Kernel_TVD (kernel.cu:2198-2661). */
int o_kernel_tvd(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b,
                 o_d2* s2, o_d2* u2, o_d3* b2, double* T, const double* T_do, int i, int method);
/* Kernel_GKR (kernel.cu:2663): the Godunov-Kolgan-Rodionov predictor-corrector
 * MHD step with its own hard-coded sides (symmetry west, outflow east, axis
 * south, inflow north), axisymmetric sources, frozen cells within r_frozen;
 * method 3 = HLLD (the reference's), 2 = HLLC-MHD. */
int o_kernel_gkr(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, o_d2* s2, o_d2* u2, o_d3* b2,
                 double* T, const double* T_do, int method);
/* its parts: POTOK_Korolkov with Q = 0 (P[8], energy in P[7]), the stencil
 * st[1..13][7] = (rho, p, vx, vy, Bx, By, Bz), the slopes DX/DY[1..5][7] and
 * the predictor of one cell (step st, radius y) */
void o_potok(double gam, double ro_L, double p_L, double v1_L, double v2_L, double v3_L, double Bx_L,
             double By_L, double Bz_L, double* P, double n1, double n2, double n3);
void o_gkr_stencil(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, int n, int m, double st[14][7]);
void o_gkr_slopes(const o_grid* g, double st[14][7], double DX[6][7], double DY[6][7]);
void o_gkr_predictor(const o_grid* g, const double* Q, const double* DX, const double* DY, double* QQ, double st,
                     double y);
/* funk_time (kernel.cu:1722-1733) */
void o_funk_time(double* T, double* T_do, double* TT, int* i);

/* ---- native path: SoA conserved planes, nfields x ny x nx ----
 * Euler fields: rho, mx, my, E.  MHD fields: rho, mx, my, Bx, By, Bz, E (Gaussian B).
 * Uout = a*Ubase + b*(Uin + dt*L(Uin)); a == 0 && b == 1 is the plain forward-Euler
 * stage and then Ubase may be NULL. When wmax != NULL the cell-centred wave-speed
 * sum max((|u|+cf)/dx + (|v|+cf)/dy) of Uout is returned through it. */
int o_stage(const o_grid* g, const o_scheme* sc, const double* Uin, const double* Ubase,
            double* Uout, double a, double b, double dt, double* wmax);
/* max over cells of the wave-speed sum of U (same formula as the stage epilogue) */
double o_wavespeed_max(const o_grid* g, const o_scheme* sc, const double* U);
/* Advance nsteps (or until t reaches t_end when t_end > 0) with the cell-centred
 * CFL step dt = cfl / wmax, clipped to t_end. integrator 1 = forward Euler,
 * 2 = SSP-RK2. tmp must hold 2 states. Returns the number of steps taken and
 * writes the final time to *t. */
long o_run(const o_grid* g, const o_scheme* sc, int integrator, double* U, double* tmp,
           long nsteps, double t_end, double* t);

/* conversions between the reference AoS primitive layout and SoA conserved */
void o_prim_to_cons(const o_grid* g, int physics, const o_d2* s, const o_d2* u, const o_d3* b, double* U);
void o_cons_to_prim(const o_grid* g, int physics, const double* U, o_d2* s, o_d2* u, o_d3* b);

/* The reference's host writers / reader restated with stdio (its ofstream
 * default format is printf "%g"). 0 on success, -1 on a bad argument or an
 * I/O failure. */
int o_write_state_text(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, const char* path);
int o_read_state_text(const o_grid* g, const char* path, o_d2* s, o_d2* u, o_d3* b);
int o_write_state_b3(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, const char* path);
int o_write_tecplot(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, int nmin,
                    const double* solution_time, const char* path);
int o_write_tecplot_derived(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, int nmin,
                            const char* path);
int o_write_line_cuts(const o_grid* g, const o_d2* s, const o_d2* u, const o_d3* b, const char* path_y0,
                      const char* path_x0);

int o_num_threads(void);
void o_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
