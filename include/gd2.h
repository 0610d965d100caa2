/*
 * gd2.h -- C ABI of libgd2.so, the MI355X (gfx950) explicit finite-volume
 * step for 2-D Euler and ideal MHD on a structured grid.
 *
 * Plain C: pointers, sizes and an opaque stream handle (a hipStream_t passed as
 * void*; NULL = the default stream). Every device pointer is caller-owned
 * device memory. Entry points are stream-ordered and asynchronous, and the
 * step, stage, flux, adjoint and layout entry points neither allocate nor
 * synchronise. The exceptions:
 *   allocate device memory: gd2_ctx_create, gd2_ensemble_create,
 *     gd2_ensemble_set_tiling and gd2_ensemble_tune (tile lists);
 *   allocate host-side HIP objects: gd2_ctx_time_kernels (events),
 *     gd2_ensemble_advance with use_graph on its first call for a step count
 *     (a capture stream and a graph);
 *   synchronise the host: gd2_ctx_time, gd2_ctx_kernel_ms, gd2_ensemble_time,
 *     gd2_ensemble_tune, gd2_ensemble_create / _set_tiling (copies of the
 *     tile list), gd2_get_diagnostics, and the destroy calls (device-wide).
 * None of these belongs inside a stream capture or a hot loop.
 *
 * Return value: GD2_OK (0) or a negative gd2_status.
 *
 * State layouts
 *   reference layout (drop-in for the reference kernels): row-major AoS
 *     primitives, index k = m*nx + n (n along x, fastest):
 *       s[k] = (rho, p), u[k] = (vx, vy), b[k] = (Bx, By, Bz) Gaussian B
 *     (kernel.cu:1738-1740, 3812-3847 of the reference).
 *   native layout: SoA conserved planes, nfields x ny x nx doubles:
 *       Euler: rho, rho*vx, rho*vy, E
 *       MHD:   rho, rho*vx, rho*vy, Bx, By, Bz, E   (Gaussian B,
 *              E = p/(gamma-1) + rho|v|^2/2 + |B|^2/(8 pi))
 */
#ifndef GD2_H
#define GD2_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y; } gd2_double2;     /* layout of double2 */
typedef struct { double x, y, z; } gd2_double3;  /* layout of double3 */

enum gd2_status {
    GD2_OK = 0,
    GD2_ERR_ARG = -1,          /* bad pointer, size or enum value */
    GD2_ERR_UNSUPPORTED = -2,  /* method / flux / boundary kind not offered by this entry point */
    GD2_ERR_HIP = -3,          /* a HIP runtime call failed (see gd2_last_hip_error) */
    GD2_ERR_NOMEM = -4,
    GD2_ERR_TOO_LARGE = -5     /* needs more than the fixed scratch of an allocation-free entry point */
};

/* Size of the fixed device scratch array of the allocation-free entry point
 * gd2_stage_vjp_noalloc (per device, held until the process ends). */
#define GD2_NOALLOC_SCRATCH_BYTES (32ll << 20)

/* Boundary kind of one side. The reference's transmissive sides copy the edge
 * cell (kernel.cu:1762-1793); its "hard" sides keep the edge cells frozen
 * (kernel.cu:1752-1758); its symmetry sides mirror with the normal velocity
 * (and normal B) negated (kernel.cu:2263-2277). Periodic has no reference
 * counterpart. FROZEN is accepted by the reference-layout entry points only.
 * The last four are the ghost rules of the reference's production MHD steps
 * (add2 / Kernel_TVD), accepted by gd2_add2 and gd2_kernel_tvd only, each on
 * its own side (add2 kernel.cu:1938-2016, Kernel_TVD kernel.cu:2227-2405):
 *   SYMMETRY (west only): mirror, negating vx, Bx and By;
 *   AXIS (south only): mirror, negating vy and Bz;
 *   OUTFLOW (east only): copy of the edge cell with vx raised to 0.01
 *     where it is negative;
 *   INFLOW (north only): rho = 1, p = 1/gamma, v = 0, and B of the edge
 *     cell for gd2_add2, B = 0 for gd2_kernel_tvd.
 * WEST_INFLOW is the west ghost of Kernel_TVD_old (kernel.cu:3323-3337,
 * 3365-3378), accepted by gd2_kernel_tvd_old on the west side only, with
 * gd2_grid.m_inf > 0: the ghost column takes the edge cell's density and
 * vy, the pressure 1/(gamma m_inf^2), and vx = -5 where the edge cell is
 * subsonic with vx > -5 (else its vx). As in the reference, the second cell
 * of a row sees that column with the edge cell's own pressure instead. */
enum gd2_bc {
    GD2_BC_TRANSMISSIVE = 0, GD2_BC_PERIODIC = 1, GD2_BC_REFLECTIVE = 2, GD2_BC_FROZEN = 3,
    GD2_BC_SYMMETRY = 4, GD2_BC_AXIS = 5, GD2_BC_OUTFLOW = 6, GD2_BC_INFLOW = 7,
    GD2_BC_WEST_INFLOW = 8
};

/* Grid and physics descriptor. The reference has all of these as
 * compile-time macros: N, M, dx, dy, x_min, y_min (kernel.cu:12-20), gamma
 * (Header.h:4-5) and the CFL factor krit (Header.h:3). */
typedef struct gd2_grid {
    int nx, ny;          /* cells along x (N) and y (M) */
    double dx, dy;       /* cell sizes */
    double x0, y0;       /* centre of cell (0,0): x_min, y_min */
    double gamma;        /* ratio of specific heats */
    double krit;         /* CFL factor of the reference-layout steps' face time step */
    int bc[4];           /* west, east, south, north: enum gd2_bc */
    /* Production options of gd2_add2 / gd2_kernel_tvd / gd2_kernel_tvd_old
     * (0 elsewhere): cells with sqrt(x^2 + y^2) < r_frozen keep their state
     * and give no time step (the inner sphere, kernel.cu:1918, 2218, 3303);
     * axisym != 0 adds the axisymmetric (y = r) source terms of
     * kernel.cu:2137-2190, 2615-2650 (MHD) and 3651-3662 (Euler), taken
     * with the signed centre y = y0 + m dy: rows below the axis are
     * accepted (Kernel_TVD and Kernel_TVD_old divide by the signed y; add2
     * evaluates them in the mirrored frame, kernel.cu:1885-2057, which is
     * the same step up to rounding); a row centred on y = 0 is GD2_ERR_ARG. */
    double r_frozen;
    int axisym;
    /* Free-stream Mach number of the WEST_INFLOW side (M_inf, Header.h:24;
     * 0 elsewhere). */
    double m_inf;
} gd2_grid;

/* ----------------------------------------------------------------------
 * Reference-layout entry points: the reference's __global__ step kernels as
 * host-callable functions with the same argument order, plus a leading grid
 * descriptor and a trailing stream. Each launches one step:
 * (s,u[,b]) -> (s2,u2[,b2]) using the time step *T_do, and folds the smallest
 * face time step krit*rad/max(|S_L|,|S_R|) of this step into *T with an exact
 * (deterministic) device-wide minimum. As in the reference, the step used is
 * the one computed in the previous step (lagged); gd2_funk_time rotates it.
 * Cartesian by default; the reference's axisymmetric 1/y source terms, inner
 * sphere and production side rules apply where the grid asks for them
 * (gd2_grid.axisym, r_frozen, side kinds 4-7; see each entry point).
 * ---------------------------------------------------------------------- */

/* Euler, first order. Replaces Ker_Dekard (kernel.cu:1736).
 * method 0 = HLL (kernel.cu:1485), 1 = HLLC (HLLC_Aleksashov, kernel.cu:853),
 * 2 = exact Riemann (Godunov_Solver_Alexashov, kernel.cu:214; the reference's
 * Ker_Dekard has no method 2). Face time step radius as in the
 * reference: dy on x-faces, dx on y-faces (kernel.cu:1798-1801). */
int gd2_ker_dekard(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u,
                   gd2_double2* s2, gd2_double2* u2, double* T, const double* T_do,
                   int method, void* stream);

/* Euler, minmod MUSCL on primitives. Replaces Kernel_TVD_old
 * (kernel.cu:3286); method as for gd2_ker_dekard. Takes r_frozen and axisym
 * (its inner sphere and 1/y terms) and its west ghost (side kind
 * WEST_INFLOW with m_inf, above). Its own sides are WEST_INFLOW, FROZEN,
 * REFLECTIVE, TRANSMISSIVE (west, east, south, north). */
int gd2_kernel_tvd_old(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u,
                       gd2_double2* s2, gd2_double2* u2, double* T, const double* T_do,
                       int method, void* stream);

/* MHD (vz = 0), first order, Powell 8-wave source. Replaces add2
 * (kernel.cu:1877); its production run sets the side kinds SYMMETRY, OUTFLOW,
 * AXIS, INFLOW, r_frozen and axisym. method 2 = HLLC-MHD, 3 = HLLD
 * (HLLDQ_Korolkov metod 2 / 3, Gas_Dinamic.cu:964). i is the step counter
 * (unused, as in the reference). */
int gd2_add2(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b,
             gd2_double2* s2, gd2_double2* u2, gd2_double3* b2, double* T, const double* T_do,
             int i, int method, void* stream);

/* MHD, minmod MUSCL on the 7 primitives + Powell source. Replaces
 * Kernel_TVD (kernel.cu:2198); production options as for gd2_add2. */
int gd2_kernel_tvd(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b,
                   gd2_double2* s2, gd2_double2* u2, gd2_double3* b2, double* T, const double* T_do,
                   int i, int method, void* stream);

/* MHD Godunov-Kolgan-Rodionov predictor-corrector step. Replaces Kernel_GKR
 * (kernel.cu:2663; predictor kernel.cu:1611, POTOK_Korolkov
 * Gas_Dinamic.cu:70): minmod slopes (the vy slot taking the pressure slope,
 * as the reference writes it), a half-step predictor of the cell and its four
 * neighbours from the physical fluxes of their half-slope face states, then
 * the update from the mean of the cell and its predictor over dt/2 with the
 * HLLD (method 3, the reference's) or HLLC-MHD (method 2) fluxes of the
 * predictor-averaged face states; axisymmetric sources throughout, no Powell
 * term. Its sides are its own: the grid must give SYMMETRY, OUTFLOW, AXIS,
 * INFLOW (west, east, south, north; else GD2_ERR_UNSUPPORTED) and axisym;
 * cells within r_frozen keep their state (2.0 in the reference). Face time
 * steps krit * rad / S with rad = dx on x-faces, dy on y-faces. */
int gd2_kernel_gkr(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b,
                   gd2_double2* s2, gd2_double2* u2, gd2_double3* b2, double* T, const double* T_do,
                   int i, int method, void* stream);

/* Time-step rotation, replaces funk_time (kernel.cu:1722):
 * *T_do = *T; *TT += *T_do; *T = 1e7; ++*i. All four are device pointers. */
int gd2_funk_time(double* T, double* T_do, double* TT, int* i, void* stream);

/* ----------------------------------------------------------------------
 * Native path: SoA conserved state, fused reconstruct->flux->update stages,
 * cell-centred CFL step computed on the device.
 * ---------------------------------------------------------------------- */
enum gd2_physics { GD2_EULER = 0, GD2_MHD = 1 };
/* Euler fluxes: HLLC = HLLC_Aleksashov (kernel.cu:853), EXACT =
 * Godunov_Solver_Alexashov (kernel.cu:214), HLL (kernel.cu:1485),
 * HLLC_KOROLKOV = HLLC_Korolkov (kernel.cu:1349: the HLL flux's wave-speed
 * bounds, star energy from the equation of state), HLLC_2D =
 * HLLC_Korolkov_2D (Gas_Dinamic.cu:265: Toro's HLLC with one-sided star
 * pressure). MHD fluxes: HLLC_MHD and HLLD = HLLDQ_Korolkov metod 2 and 3
 * (Gas_Dinamic.cu:964). */
enum gd2_flux { GD2_FLUX_HLLC = 1, GD2_FLUX_EXACT = 2, GD2_FLUX_HLLC_MHD = 3, GD2_FLUX_HLLD = 4, GD2_FLUX_HLL = 5,
                GD2_FLUX_HLLC_KOROLKOV = 6, GD2_FLUX_HLLC_2D = 7 };
enum gd2_integrator { GD2_FORWARD_EULER = 1, GD2_SSPRK2 = 2 };

typedef struct gd2_scheme {
    int physics;      /* enum gd2_physics */
    int flux;         /* enum gd2_flux (HLLC/EXACT/HLL for Euler, HLLC_MHD/HLLD for MHD) */
    int order;        /* 1 = piecewise constant, 2 = minmod MUSCL on primitives */
    int integrator;   /* enum gd2_integrator */
    double cfl;       /* dt = cfl / max_cells((|u|+c)/dx + (|v|+c)/dy) */
} gd2_scheme;

/* Number of conserved fields of a physics kind (4 or 7). */
int gd2_nfields(int physics);

/* One stage on caller-owned SoA device buffers:
 *   Uout = a*Ubase + b*(Uin - dt/(dx dy) * sum of face fluxes(Uin))
 * (a = 0, b = 1 is a forward-Euler step; Ubase may then be NULL; Uout may
 * alias Ubase but not Uin). dt is taken from *dt_dev when dt_dev != NULL,
 * else from dt. When wmax_bits != NULL (device, 8 bytes) the cell-centred
 * wave-speed bound of Uout is max-folded into it as the bit pattern of a
 * non-negative double. Clamp hits (rho <= 0, p <= 0) are counted in the
 * library's device diagnostics. */
int gd2_stage(const gd2_grid* g, const gd2_scheme* sc, const double* Uin, const double* Ubase,
              double* Uout, double a, double b, double dt, const double* dt_dev,
              unsigned long long* wmax_bits, void* stream);

/* One SSP-RK2 step of the Euler scheme in a single pass, out of place:
 *   U1 = Uin + dt L(Uin),  Uout = Uin/2 + (U1 + dt L(U1))/2
 * with the intermediate state U1 kept on chip. Bitwise equal to
 *   gd2_stage(Uin -> S, a=0, b=1); gd2_stage(S, Ubase=Uin -> Uout, a=b=1/2)
 * (including the wave-speed fold into wmax_bits and the clamp counts).
 * Ustage may be null; otherwise it receives U1 (bitwise S above), e.g. for
 * an adjoint sweep that needs the stage state. Uout and Ustage must not alias
 * Uin or each other. Euler physics (HLLC or exact flux) only; other schemes
 * return GD2_ERR_UNSUPPORTED. No reference counterpart (the reference steps
 * with forward Euler); used by gd2_ctx_advance for SSP-RK2 Euler. */
int gd2_ssprk2_step(const gd2_grid* g, const gd2_scheme* sc, const double* Uin, double* Uout,
                    double* Ustage, double dt, const double* dt_dev, unsigned long long* wmax_bits,
                    void* stream);

/* How SSP-RK2 Euler steps are taken where the library chooses (the solver
 * context): GD2_STEP_AUTO uses gd2_ssprk2_step on grids of at least 2^21
 * cells and two gd2_stage launches on smaller ones (the faster of the two on
 * MI355X); the others force one path. Both paths give the same bits.
 * Process-wide; returns the previous mode. No reference counterpart. */
enum gd2_step_path { GD2_STEP_AUTO = 0, GD2_STEP_TWO_PASS = 1, GD2_STEP_ONE_PASS = 2 };
int gd2_set_step_path(int mode);
/* 1 when a step of this grid and scheme goes through gd2_ssprk2_step under
 * the current mode, 0 when through two gd2_stage launches. */
int gd2_one_pass_preferred(const gd2_grid* g, const gd2_scheme* sc);

/* max over cells of (|u|+c)/dx + (|v|+c)/dy of U, folded into *wmax_bits
 * (device) as above; the caller zeroes *wmax_bits first. */
int gd2_wavespeed_max(const gd2_grid* g, const gd2_scheme* sc, const double* U,
                      unsigned long long* wmax_bits, void* stream);

/* Solver context: owns one scratch state and a device clock. */
typedef struct gd2_ctx gd2_ctx;

int gd2_ctx_create(const gd2_grid* g, const gd2_scheme* sc, gd2_ctx** out);
int gd2_ctx_destroy(gd2_ctx* ctx);
/* Set the clock to time t0 and compute the first dt from U (device). */
int gd2_ctx_start(gd2_ctx* ctx, const double* U, double t0, void* stream);
/* Advance the device state U in place by nsteps steps. Each step takes
 * dt = min(cfl / wmax(U^n), t_end - t) when t_end > 0, else cfl / wmax(U^n);
 * steps requested after t reached t_end are exact no-ops. No host sync. A
 * one-pass SSP-RK2 step with t_end <= 0 is a single kernel launch that also
 * updates the device clock; otherwise a one-thread clock kernel precedes
 * each step's kernels. */
int gd2_ctx_advance(gd2_ctx* ctx, double* U, int nsteps, double t_end, void* stream);
/* Synchronises the stream and reads the clock: time, steps taken, last dt. */
int gd2_ctx_time(gd2_ctx* ctx, double* t, long long* steps, double* dt, void* stream);

/* ----------------------------------------------------------------------
 * Ensembles: n independent Euler problems (SSP-RK2, each with its own grid,
 * boundary kinds and CFL step) advanced together. gd2_ensemble_advance runs
 * up to 64 steps in one persistent launch for the whole ensemble: resident
 * waves take (step, tile) entries of device work lists in step order (one
 * list per XCD when the problems can be spread evenly over the XCDs), and a
 * tile of a problem's step waits until that problem's previous step is
 * complete (per-problem clocks and completion counters on the device; a wait
 * that exceeds its spin limit stops waiting and is reported as GD2_ERR_HIP by
 * the next gd2_ensemble_time, as is a launch that left tiles untaken). An odd
 * step count takes its first step as two stage launches. gd2_ensemble_advance can capture its launches in a hipGraph
 * and replay it. No reference counterpart (the reference steps a single
 * grid); results are bitwise those of running each problem alone through
 * gd2_ctx_advance. Launches on one ensemble must not overlap (they share its
 * work list and second state buffers).
 * ---------------------------------------------------------------------- */
typedef struct gd2_ensemble gd2_ensemble;

/* U[p] is problem p's caller-owned SoA conserved state (4 planes, nx, ny >=
 * 2). Allocates a second state for every problem, the device descriptors and
 * the work lists. */
int gd2_ensemble_create(int n, const gd2_grid* grids, const gd2_scheme* sc, double* const* U,
                        gd2_ensemble** out);
int gd2_ensemble_destroy(gd2_ensemble* e);
/* Strip height of the persistent launch's tiles (rows per wave, 0 = default,
 * at least 2) and its number of waves (0 = as many as fit on the device). */
int gd2_ensemble_set_tiling(gd2_ensemble* e, int rows, int blocks);
/* The strip height and wave count in use, and the tiles per step. */
int gd2_ensemble_get_tiling(gd2_ensemble* e, int* rows, int* blocks, int* ntiles);
/* Reset every clock to t = 0 and compute each problem's first dt. */
int gd2_ensemble_start(gd2_ensemble* e, void* stream);
/* Advance every problem by nsteps; use_graph != 0 replays a captured hipGraph
 * of the nsteps steps (captured on first use, cached). */
int gd2_ensemble_advance(gd2_ensemble* e, int nsteps, int use_graph, void* stream);
/* Synchronises the stream; reads problem p's clock. GD2_ERR_HIP if a
 * persistent launch since the previous call hit its wait limit. */
int gd2_ensemble_time(gd2_ensemble* e, int p, double* t, long long* steps, double* dt, void* stream);
/* Times dry persistent launches (they read the states and write only the
 * library's second buffers; clocks untouched) over candidate strip heights
 * and wave counts and keeps the fastest. Call after gd2_ensemble_start.
 * Synchronises the stream; best_step_ms is the best time per step. */
int gd2_ensemble_tune(gd2_ensemble* e, void* stream, int* rows, int* blocks, double* best_step_ms);

/* Record each taken step's dt into dt_log[step] (device, capacity entries);
 * capacity 0 switches the log off. */
int gd2_ctx_log_dt(gd2_ctx* ctx, double* dt_log, long long capacity);

/* Kernel timing (no reference counterpart; the reference times whole phases
 * with events around its loop, kernel.cu:4087-4091). Not while the stream is
 * captured: gd2_ctx_advance then records no events. gd2_ctx_advance records
 * a pair of HIP events around the kernel launches of every stride-th step
 * (stride >= 1; the first step after this call included) until `capacity`
 * steps are timed: the one-pass step kernel, or the step's stage kernels, not
 * its one-thread clock kernel. capacity 0 switches it off. Allocates
 * host-side events only. */
int gd2_ctx_time_kernels(gd2_ctx* ctx, int capacity, int stride);
/* Mean duration in ms of the steps timed since the last call (waits for the
 * last of them: synchronises with the stream), their number in *count (may
 * be NULL); starts the next count. */
int gd2_ctx_kernel_ms(gd2_ctx* ctx, double* mean_ms, int* count);

/* ----------------------------------------------------------------------
 * Discrete adjoint of the stage operator: Euler with the HLLC flux, or MHD
 * with the HLLC-MHD or HLLD flux (Powell term included); order 1 or 2.
 * No reference counterpart. For L(U) = -(1/(dx dy)) * (sum of outward face
 * fluxes times face lengths [+ Powell source]), computes
 *     W = a*A + b*B + c * J_L(U)^T V
 * (A, B may be NULL; W may alias A or B but not U or V). The minmod branch of
 * every slope is the one the forward stage takes on U (weights 1/2-1/2 on a
 * tie, 0 on the zero branch); the rho/p <= 0 face fall-back and the boundary
 * ghost fill are differentiated as copies; the clamps are not differentiated.
 * workspace: device memory of gd2_vjp_workspace_bytes(g, sc) bytes. One
 * SSP-RK2 step adjoint, with U1 = U + dt L(U):
 *     lam1 = lam/2 + (dt/2) J_L(U1)^T lam;
 *     lam0 = lam/2 + lam1 + dt J_L(U)^T lam1.
 * ---------------------------------------------------------------------- */
long long gd2_vjp_workspace_bytes(const gd2_grid* g, const gd2_scheme* sc);
int gd2_stage_vjp(const gd2_grid* g, const gd2_scheme* sc, const double* U, const double* V,
                  const double* A, double a, const double* B, double b, double c, double* W,
                  void* workspace, void* stream);

/* gd2_stage_vjp without the workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. It calls
 * no allocating routine: the workspace is a fixed device array of
 * GD2_NOALLOC_SCRATCH_BYTES bytes inside the library (one per device, held
 * until the process ends), looked up on the first call on each device.
 * Returns GD2_ERR_TOO_LARGE when gd2_vjp_workspace_bytes(g, sc) exceeds it:
 * Euler up to max(nx, ny) = 131072, MHD up to about 200 000 cells. Every call
 * on one device shares that device's array, so calls that overlap on one
 * device (concurrent streams or host threads) race; calls in one stream are
 * safe, since each call writes every entry it reads. Results are those of
 * gd2_stage_vjp. */
int gd2_stage_vjp_noalloc(const gd2_grid* g, const gd2_scheme* sc, const double* U, const double* V,
                          const double* A, double a, const double* B, double b, double c, double* W,
                          void* stream);

/* ----------------------------------------------------------------------
 * The reference's interface fluxes evaluated pointwise: one GPU thread per
 * face calls the flux function with the reference's own lab-frame arguments
 * and outputs (its device-function flux interface, Header.h:41-65 and the
 * flux functions of kernel.cu / Gas_Dinamic.cu). Arrays of n entries in
 * device memory; rad is the face size passed as `rad`, and dt[i] receives the
 * function's return value, the face time step krit * rad / S.
 * ---------------------------------------------------------------------- */
enum gd2_face_flux {
    GD2_FACE_HLL = 0,               /* HLL, kernel.cu:1485 */
    GD2_FACE_HLLC_ALEKSASHOV = 1,   /* HLLC_Aleksashov, kernel.cu:853 */
    GD2_FACE_GODUNOV = 2,           /* Godunov_Solver_Alexashov (exact), kernel.cu:214 */
    GD2_FACE_HLLC_ALEKSASHOV2 = 3,  /* HLLC_Aleksashov2, kernel.cu:1167 */
    GD2_FACE_HLLC_KOROLKOV = 4,     /* HLLC_Korolkov, kernel.cu:1349 */
    GD2_FACE_HLLC_KOROLKOV_2D = 5,  /* HLLC_Korolkov_2D, Gas_Dinamic.cu:265 */
    GD2_FACE_HLLCQ_KOROLKOV_2D = 6, /* HLLCQ_Korolkov_2D, Gas_Dinamic.cu:359 (passive scalar) */
    GD2_FACE_HLLCQ_ALEKSASHOV = 7   /* HLLCQ_Aleksashov with B = 0 as its callers pass it, Gas_Dinamic.cu:465 */
};

/* Euler: Ls[i] = (rho, p), Lu[i] = (vx, vy) on the left, Rs, Ru on the right,
 * normal[i] = (n1, n2); out Ps[i] = (F_rho, F_E), Pu[i] = (F_mx, F_my) along
 * the normal. QL, QR, PQ: the passive scalar and its flux, for the two Q
 * kinds only (NULL otherwise). */
int gd2_euler_face_fluxes(int kind, double gamma, double krit, double rad, long long n,
                          const gd2_double2* Ls, const gd2_double2* Lu, const gd2_double2* Rs,
                          const gd2_double2* Ru, const double* QL, const double* QR,
                          const gd2_double2* normal, gd2_double2* Ps, gd2_double2* Pu, double* PQ,
                          double* dt, void* stream);
/* MHD, HLLDQ_Korolkov (Gas_Dinamic.cu:964) metod 2 (HLLC-MHD) or 3 (HLLD)
 * without its passive scalar: L, R of n x 8 doubles (rho, p, vx, vy, vz, Bx,
 * By, Bz; Gaussian B), normal of n x 3 (its frame table, Gas_Dinamic.cu:
 * 978-1019, covers the six axis directions); out P of n x 8 (F_rho, F_mx,
 * F_my, F_mz, F_Bx, F_By, F_Bz, F_E). */
int gd2_mhd_face_fluxes(int metod, double gamma, double krit, double rad, long long n, const double* L,
                        const double* R, const double* normal, double* P, double* dt, void* stream);

/* Layout conversion between the reference AoS primitives and SoA conserved. */
int gd2_prim_to_cons(const gd2_grid* g, int physics, const gd2_double2* s, const gd2_double2* u,
                     const gd2_double3* b, double* U, void* stream);
int gd2_cons_to_prim(const gd2_grid* g, int physics, const double* U, gd2_double2* s,
                     gd2_double2* u, gd2_double3* b, void* stream);

/* Device diagnostics: clamp hits (rho, p) and exact-solver iteration-cap hits
 * (newton, bisection). Synchronises the device. */
typedef struct gd2_diag {
    long long clamp_rho, clamp_p, newton_cap_hits, bisect_cap_hits;
} gd2_diag;
int gd2_get_diagnostics(gd2_diag* out, int reset);

/* ----------------------------------------------------------------------
 * On-disk formats of the reference's host driver. HOST buffers in the
 * reference layout (the reference copies its state to host_s / host_u /
 * host_b before writing); no device work. Numbers are written as its
 * std::ofstream writes them: printf "%g" (6 significant digits). Cell
 * centres are x = x0 + n*dx, y = y0 + m*dy. GD2_ERR_ARG on a bad argument or
 * an I/O failure. GD2_IO_THREADS caps the formatting threads (default: the
 * host's, at most 16).
 * ---------------------------------------------------------------------- */

/* Text state, one cell per line: "x y rho p vx vy Bx By Bz"
 * (instable_for_inner_b3.txt, kernel.cu:4257-4260); b may be NULL (zeros). */
int gd2_write_state_text(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u,
                         const gd2_double3* b, const char* path);
/* Restart reader (kernel.cu:3944-3969): 9 whitespace-separated numbers per
 * cell, the first two (coordinates) skipped; b may be NULL (Euler). */
int gd2_read_state_text(const gd2_grid* g, const char* path, gd2_double2* s, gd2_double2* u,
                        gd2_double3* b);
/* save_b3.dat (kernel.cu:4261-4267): 7 native-endian doubles per cell,
 * x, y, rho, p, vx, vy, Bz. */
int gd2_write_state_b3(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u,
                       const gd2_double3* b, const char* path);
/* Tecplot FEPOINT quadrilateral zone of every nmin-th cell in x and y
 * (print_file_mini2 kernel.cu:3724-3808 with b, print_file_mini
 * kernel.cu:3682-3722 with b == NULL; the reference's Nmin is 3):
 * X Y Ro P Vx Vy [Bx By Bz] Max T, Max = |v| / sqrt(gamma p / rho),
 * T = p / rho, on that writer's own coordinates (cell centres stretched
 * over [x0, right edge] with an (nx - 1) spacing); solution_time != NULL adds
 * the SOLUTIONTIME clause. */
int gd2_write_tecplot(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u,
                      const gd2_double3* b, int nmin, const double* solution_time, const char* path);

/* The end-of-run derived Tecplot file param_for_texplot_mini.txt
 * (kernel.cu:4231-4233, 4384-4460): every nmin-th cell on the cell centres,
 * X Y Ro P Vx Vy Bx By Bz Max T PP Fx Fy zav with PP = p + Bz^2/8pi,
 * Fx = Bz dBz/dx / 4pi, Fy = (-Bz^2/y - Bz dBz/dy) / 4pi (central
 * differences, forward on the first column / row, 0 on the last) and the
 * vorticity zav (0 on the first and last column / row). */
int gd2_write_tecplot_derived(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u,
                              const gd2_double3* b, int nmin, const char* path);
/* The end-of-run line cuts (kernel.cu:4238-4246, 4283-4374): param_y=0.txt,
 * the first row's cells 5 .. nx-2 (X Ro P Vx Bz Max T PP Fx Fy dP), and
 * param_x=0.txt, the first column's cells 5 .. ny-2 (Y Ro P Vy Bz Max T PP Fx
 * Fy dP), with forward y differences and dP the forward difference of PP;
 * either path may be NULL. */
int gd2_write_line_cuts(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u,
                        const gd2_double3* b, const char* path_y0, const char* path_x0);

/* Last HIP error code seen by the library (hipError_t value). */
int gd2_last_hip_error(void);
/* Library version string. */
const char* gd2_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GD2_H */
