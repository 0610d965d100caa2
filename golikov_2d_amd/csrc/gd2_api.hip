// gd2_api.hip -- the C ABI of libgd2.so (declared in include/gd2.h):
// argument checks, launch configuration, the device clock of the native
// solver context and the device diagnostics. All compute is in the kernels of
// gd2_euler.hpp / gd2_mhd.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <atomic>
#include <mutex>
#include <map>
#include <new>
#include <vector>

#include "../../include/gd2.h"
#include "gd2_euler.hpp"
#include "gd2_launch.hpp"
#include "gd2_rk2.hpp"
#include "gd2_adjoint.hpp"
#include "gd2_mhd_launch.hpp"
#include "gd2_ensemble.hpp"
#include "gd2_adjoint.hpp"

namespace gd2 {

__device__ DeviceDiag g_diag;

// Fixed device scratch of the allocation-free adjoint entry point
// (gd2_stage_vjp_noalloc): GD2_NOALLOC_SCRATCH_BYTES per device, held by
// every device the library runs on until the process ends. Host code never
// passes this name to a kernel (on the host it is a stand-in, not a device
// address); it uses the address hipGetSymbolAddress returns for the current
// device (noalloc_scratch below).
static __device__ __align__(256) double scratch[GD2_NOALLOC_SCRATCH_BYTES / sizeof(double)];

static thread_local int g_last_hip = 0;

static int hip_fail(hipError_t e) {
    g_last_hip = (int)e;
    return GD2_ERR_HIP;
}

#define GD2_HIP(call)                                 \
    do {                                              \
        hipError_t e_ = (call);                       \
        if (e_ != hipSuccess) return hip_fail(e_);    \
    } while (0)

// Device address of `scratch` on the current device, looked up once per
// device with that device current and kept per device (each device holds its
// own copy of the array at its own address). Null when the lookup fails.
static void* noalloc_scratch() {
    static std::mutex mu;
    static std::map<int, void*> addr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        hip_fail(e);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    const auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol);
        return nullptr;
    }
    addr[dev] = p;
    return p;
}

static DeviceDiag* diag_ptr() {
    static DeviceDiag* p = nullptr;
    if (!p) {
        void* q = nullptr;
        if (hipGetSymbolAddress(&q, HIP_SYMBOL(g_diag)) != hipSuccess) return nullptr;
        p = static_cast<DeviceDiag*>(q);
    }
    return p;
}

// prod: which production options of the reference-layout steps an entry
// point takes: kProdSides (side kinds SYMMETRY / AXIS / OUTFLOW / INFLOW,
// each on its own side: the MHD steps), kProdOptions (the frozen sphere
// and the axisymmetric sources: the MHD steps and Kernel_TVD_old) and
// kProdWestInflow (the WEST_INFLOW side: Kernel_TVD_old); everywhere else
// they are rejected rather than ignored.
enum { kProdSides = 1, kProdOptions = 2, kProdAll = 3, kProdWestInflow = 4 };
static bool grid_ok(const gd2_grid* g, bool allow_frozen, int prod = 0) {
    if (!g || g->nx < 1 || g->ny < 1) return false;
    if (!(g->dx > 0.0) || !(g->dy > 0.0) || !(g->gamma > 1.0)) return false;
    if ((long long)g->nx * g->ny > (1ll << 40)) return false;
    static const int own_side[4] = {GD2_BC_SYMMETRY, GD2_BC_OUTFLOW, GD2_BC_AXIS, GD2_BC_INFLOW};
    for (int s = 0; s < 4; ++s) {
        if (g->bc[s] == GD2_BC_WEST_INFLOW) {
            if (s != 0 || !(prod & kProdWestInflow) || !std::isfinite(g->m_inf) || !(g->m_inf > 0.0)) return false;
            continue;
        }
        if (g->bc[s] < GD2_BC_TRANSMISSIVE || g->bc[s] > GD2_BC_INFLOW) return false;
        if (g->bc[s] == GD2_BC_FROZEN && !allow_frozen) return false;
        if (g->bc[s] >= GD2_BC_SYMMETRY && (!(prod & kProdSides) || g->bc[s] != own_side[s])) return false;
    }
    if (!(prod & kProdOptions) && (g->r_frozen != 0.0 || g->axisym != 0)) return false;
    if (!(g->r_frozen >= 0.0)) return false;
    // the 1/y terms take the signed centre y0 + m dy (rows below the axis
    // included, as Kernel_TVD and Kernel_TVD_old do and as add2's mirrored
    // evaluation of such rows amounts to); a row centred on the axis would
    // divide by zero
    if (g->axisym) {
        if (!std::isfinite(g->y0)) return false;
        const double k = std::nearbyint(-g->y0 / g->dy);
        if (k >= 0.0 && k < (double)g->ny && std::fabs(g->y0 + k * g->dy) <= 1e-9 * g->dy) return false;
    }
    // a periodic side needs a periodic opposite side
    if ((g->bc[0] == GD2_BC_PERIODIC) != (g->bc[1] == GD2_BC_PERIODIC)) return false;
    if ((g->bc[2] == GD2_BC_PERIODIC) != (g->bc[3] == GD2_BC_PERIODIC)) return false;
    return true;
}

int hip_error_code(hipError_t e) { return e == hipSuccess ? GD2_OK : hip_fail(e); }

DeviceDiag* device_diag() { return diag_ptr(); }

// Rows per wave strip. Default: long enough to amortise the 4-row y-halo,
// short enough to give every CU several waves; GD2_TILE_ROWS overrides it.
int tile_rows_for(int nx, int ny) {
    static int env = -1;
    if (env < 0) {
        const char* e = std::getenv("GD2_TILE_ROWS");
        env = e ? std::atoi(e) : 0;
    }
    if (env > 0) return env;
    const long long tiles_x = (nx + kTileW - 1) / kTileW;
    int rows = 64;
    while (rows > 8 && tiles_x * ((ny + rows - 1) / rows) < 16384) rows /= 2;
    return rows;
}

// Rows per strip for a launch of one-wave strips of a kernel: waves run in
// rounds of (CUs x resident waves per CU); a round takes about as long as
// one strip (its rows plus the rows re-read around it, `overhead`), so pick
// the height minimising rounds x (rows + overhead). GD2_TILE_ROWS overrides.
int strip_rows(const void* kernel, int tiles_x, int ny, int overhead, int max_rows, int min_rows) {
    static int env = -1;
    if (env < 0) {
        const char* e = std::getenv("GD2_TILE_ROWS");
        env = e ? std::atoi(e) : 0;
    }
    if (env > 0) return env;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    struct Entry {
        const void* k;
        int dev, cus, per_cu;
    };
    static Entry cache[64];
    static int ncache = 0;
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    for (int i = 0; i < ncache; ++i)
        if (cache[i].k == kernel && cache[i].dev == dev) {
            cus = cache[i].cus;
            per_cu = cache[i].per_cu;
        }
    if (cus == 0) {
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kWave, 0) != hipSuccess || per_cu <= 0)
            per_cu = 4;
        if (ncache < 64) cache[ncache++] = Entry{kernel, dev, cus, per_cu};
    }
    const long long slots = (long long)cus * per_cu;
    int best = 16;
    long long best_cost = -1;
    for (int rows = min_rows; rows <= max_rows; ++rows) {
        const long long waves = (long long)tiles_x * ((ny + rows - 1) / rows);
        // strips under four rows only where every wave has a SIMD of its own
        // (a lone wave's march is then the critical path); with shared SIMDs
        // their extra lead-in and halo rows cost more than they save
        if (rows < 4 && waves > (long long)cus * 4) continue;
        const long long cost = ((waves + slots - 1) / slots) * (rows + overhead);
        if (best_cost < 0 || cost <= best_cost) {  // ties: the taller strip (less halo)
            best_cost = cost;
            best = rows;
        }
    }
    return best;
}

static void fill_common(EulerArgs& A, const gd2_grid* g) {
    std::memset(&A, 0, sizeof A);
    A.nx = g->nx;
    A.ny = g->ny;
    A.plane = (long long)g->nx * g->ny;
    A.dx = g->dx;
    A.dy = g->dy;
    A.idx = 1.0 / g->dx;
    A.idy = 1.0 / g->dy;
    A.gamma = g->gamma;
    A.g1 = g->gamma - 1.0;
    A.inv_g1 = 1.0 / (g->gamma - 1.0);
    A.krit = g->krit;
    for (int s = 0; s < 4; ++s) A.bc[s] = g->bc[s];
    A.diag = diag_ptr();
    A.rows = tile_rows_for(g->nx, g->ny);
    A.tiles_x = (g->nx + kTileW - 1) / kTileW;
}

template <int LAYOUT, bool W, bool BASE>
static int dispatch_euler(int order, int flux, const EulerArgs& A, hipStream_t st) {
    if (order == 2) {
        if (flux == GD2_FLUX_HLLC) return launch_euler<2, GD2_FLUX_HLLC, LAYOUT, W, BASE>(A, st);
        if (flux == GD2_FLUX_EXACT) return launch_euler<2, GD2_FLUX_EXACT, LAYOUT, W, BASE>(A, st);
        if (flux == GD2_FLUX_HLL) return launch_euler<2, GD2_FLUX_HLL, LAYOUT, W, BASE>(A, st);
    } else if (order == 1) {
        if (flux == GD2_FLUX_HLLC) return launch_euler<1, GD2_FLUX_HLLC, LAYOUT, W, BASE>(A, st);
        if (flux == GD2_FLUX_EXACT) return launch_euler<1, GD2_FLUX_EXACT, LAYOUT, W, BASE>(A, st);
        if (flux == GD2_FLUX_HLL) return launch_euler<1, GD2_FLUX_HLL, LAYOUT, W, BASE>(A, st);
    }
    if (is_hllc_variant(flux)) return launch_euler_variant(order, flux, LAYOUT, W, BASE, kRefPlain, A, st);
    return GD2_ERR_UNSUPPORTED;
}

template <bool KEEP>
static int dispatch_rk2(int order, int flux, const EulerArgs& A, hipStream_t st) {
    if (order == 2) {
        if (flux == GD2_FLUX_HLLC) return launch_rk2<2, GD2_FLUX_HLLC, KEEP>(A, st);
        if (flux == GD2_FLUX_EXACT) return launch_rk2<2, GD2_FLUX_EXACT, KEEP>(A, st);
        if (flux == GD2_FLUX_HLL) return launch_rk2<2, GD2_FLUX_HLL, KEEP>(A, st);
    } else if (order == 1) {
        if (flux == GD2_FLUX_HLLC) return launch_rk2<1, GD2_FLUX_HLLC, KEEP>(A, st);
        if (flux == GD2_FLUX_EXACT) return launch_rk2<1, GD2_FLUX_EXACT, KEEP>(A, st);
        if (flux == GD2_FLUX_HLL) return launch_rk2<1, GD2_FLUX_HLL, KEEP>(A, st);
    }
    if (is_hllc_variant(flux)) return launch_rk2_variant(order, flux, KEEP, A, st);
    return GD2_ERR_UNSUPPORTED;
}

static bool scheme_ok(const gd2_scheme* sc) {
    if (!sc || (sc->order != 1 && sc->order != 2)) return false;
    if (sc->physics == GD2_EULER)
        return sc->flux == GD2_FLUX_HLLC || sc->flux == GD2_FLUX_EXACT || sc->flux == GD2_FLUX_HLL ||
               is_hllc_variant(sc->flux);
    if (sc->physics == GD2_MHD) return sc->flux == GD2_FLUX_HLLC_MHD || sc->flux == GD2_FLUX_HLLD;
    return false;
}

// ---------------------------------------------------------------- clock
__global__ void clock_begin_kernel(Clock* c, int cur, double t_end, double* dt_log, long long cap) {
    const long long n = c->step;
    clock_begin(c, cur, t_end);
    if (dt_log && c->step > n && n < cap) dt_log[n] = c->dt;  // dt of step n (taken steps only)
}

__global__ void clock_start_kernel(Clock* c, double t0, double cfl) {
    c->t = t0;
    c->dt = 0.0;
    c->cfl = cfl;
    c->w[0] = 0ull;
    c->w[1] = 0ull;
    c->w[2] = 0ull;
    c->step = 0;
}

__global__ void funk_time_kernel(double* T, double* T_do, double* TT, int* i) {
    // kernel.cu:1722-1733 (its every-5000-steps printf is left out)
    *T_do = *T;
    *TT = *TT + *T_do;
    *T = 10000000;
    *i = *i + 1;
}

__global__ void prim_to_cons_kernel(const double2* s, const double2* u, const double3* b, double* U, long long n,
                                    double gamma, int mhd) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        const double r = s[k].x, p = s[k].y, vx = u[k].x, vy = u[k].y;
        U[k] = r;
        U[n + k] = r * vx;
        U[2 * n + k] = r * vy;
        if (mhd) {
            const double3 B = b[k];
            U[3 * n + k] = B.x;
            U[4 * n + k] = B.y;
            U[5 * n + k] = B.z;
            U[6 * n + k] = p / (gamma - 1.0) + 0.5 * r * (vx * vx + vy * vy) + (B.x * B.x + B.y * B.y + B.z * B.z) / kCpi8;
        } else {
            U[3 * n + k] = p / (gamma - 1.0) + 0.5 * r * (vx * vx + vy * vy);
        }
    }
}

__global__ void cons_to_prim_kernel(const double* U, double2* s, double2* u, double3* b, long long n, double gamma,
                                    int mhd) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        const double r = U[k], mx = U[n + k], my = U[2 * n + k];
        const double inv = 1.0 / r;
        const double vx = mx * inv, vy = my * inv;
        const double ke = 0.5 * (mx * vx + my * vy);
        double p;
        if (mhd) {
            const double bx = U[3 * n + k], by = U[4 * n + k], bz = U[5 * n + k];
            p = (gamma - 1.0) * (U[6 * n + k] - ke - (bx * bx + by * by + bz * bz) / kCpi8);
            if (b) b[k] = make_double3(bx, by, bz);
        } else {
            p = (gamma - 1.0) * (U[3 * n + k] - ke);
        }
        s[k] = make_double2(r, p);
        u[k] = make_double2(vx, vy);
    }
}

static unsigned ew_blocks(long long n) {
    long long b = (n + 255) / 256;
    if (b > 2048) b = 2048;
    if (b < 1) b = 1;
    return (unsigned)b;
}



}  // namespace gd2

using namespace gd2;

struct gd2_ctx {
    gd2_grid g;
    gd2_scheme sc;
    double* scratch;  // one state (nfields planes)
    Clock* clock;
    long long host_steps;  // steps enqueued (parity of the wave-speed slots)
    double* dt_log;        // optional device log of each step's dt
    long long dt_cap;
    // optional per-step kernel timing: a pair of events around each step's
    // kernel launches, for the first ev_cap steps after gd2_ctx_time_kernels
    hipEvent_t* ev = nullptr;
    int ev_cap = 0, ev_n = 0, ev_stride = 1;
    long long ev_seen = 0;  // steps enqueued since timing was switched on
};

static void drop_events(gd2_ctx* c) {
    for (int i = 0; i < 2 * c->ev_cap; ++i) (void)hipEventDestroy(c->ev[i]);
    delete[] c->ev;
    c->ev = nullptr;
    c->ev_cap = c->ev_n = 0;
}

extern "C" {

int gd2_nfields(int physics) { return physics == GD2_MHD ? 7 : (physics == GD2_EULER ? 4 : GD2_ERR_ARG); }

int gd2_last_hip_error(void) { return g_last_hip; }

const char* gd2_version(void) { return "gd2 0.1.0 (gfx950)"; }

int gd2_stage(const gd2_grid* g, const gd2_scheme* sc, const double* Uin, const double* Ubase, double* Uout,
              double a, double b, double dt, const double* dt_dev, unsigned long long* wmax_bits, void* stream) {
    if (!grid_ok(g, false) || !scheme_ok(sc) || !Uin || !Uout || Uin == Uout) return GD2_ERR_ARG;
    const bool plain = (a == 0.0 && b == 1.0);
    if (!plain && !Ubase) return GD2_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (sc->physics == GD2_MHD) return mhd_stage(g, sc, Uin, Ubase, Uout, a, b, dt, dt_dev, wmax_bits, diag_ptr(), st);
    EulerArgs A;
    fill_common(A, g);
    A.Uin = Uin;
    A.Ubase = Ubase;
    A.Uout = Uout;
    A.a = a;
    A.b = b;
    A.dt = dt;
    A.dt_dev = dt_dev;
    A.wmax_bits = wmax_bits;
    if (!A.diag) return GD2_ERR_HIP;
    if (plain) {
        if (wmax_bits) return dispatch_euler<kSoA, true, false>(sc->order, sc->flux, A, st);
        return dispatch_euler<kSoA, false, false>(sc->order, sc->flux, A, st);
    }
    if (wmax_bits) return dispatch_euler<kSoA, true, true>(sc->order, sc->flux, A, st);
    return dispatch_euler<kSoA, false, true>(sc->order, sc->flux, A, st);
}

// One-pass SSP-RK2 step; with clk != null the kernel takes its dt from the
// clock (slot clk_slot) and one of its threads does the clock update.
static int ssprk2_launch(const gd2_grid* g, const gd2_scheme* sc, const double* Uin, double* Uout, double* Ustage,
                         double dt, const double* dt_dev, unsigned long long* wmax_bits, Clock* clk, int clk_slot,
                         double* dt_log, long long dt_cap, hipStream_t st) {
    if (!grid_ok(g, false) || !scheme_ok(sc) || !Uin || !Uout || Uin == Uout) return GD2_ERR_ARG;
    if (Ustage && (Ustage == Uin || Ustage == Uout)) return GD2_ERR_ARG;
    if (sc->physics != GD2_EULER) return GD2_ERR_UNSUPPORTED;
    // the boundary rows and columns of the intermediate state are built from
    // its first and last two rows / columns
    if (g->nx < 2 || g->ny < 2) return GD2_ERR_UNSUPPORTED;
    EulerArgs A;
    fill_common(A, g);
    A.Uin = Uin;
    A.Uout = Uout;
    A.Ustage = Ustage;
    A.a = 0.5;
    A.b = 0.5;
    A.dt = dt;
    A.dt_dev = dt_dev;
    A.wmax_bits = wmax_bits;
    A.clk = clk;
    A.clk_slot = clk_slot;
    A.dt_log = dt_log;
    A.dt_cap = dt_cap;
    if (!A.diag) return GD2_ERR_HIP;
    if (Ustage) return dispatch_rk2<true>(sc->order, sc->flux, A, st);
    return dispatch_rk2<false>(sc->order, sc->flux, A, st);
}

int gd2_ssprk2_step(const gd2_grid* g, const gd2_scheme* sc, const double* Uin, double* Uout, double* Ustage,
                    double dt, const double* dt_dev, unsigned long long* wmax_bits, void* stream) {
    return ssprk2_launch(g, sc, Uin, Uout, Ustage, dt, dt_dev, wmax_bits, nullptr, 0, nullptr, 0,
                         (hipStream_t)stream);
}

int gd2_wavespeed_max(const gd2_grid* g, const gd2_scheme* sc, const double* U, unsigned long long* wmax_bits,
                      void* stream) {
    if (!grid_ok(g, true) || !scheme_ok(sc) || !U || !wmax_bits) return GD2_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const long long n = (long long)g->nx * g->ny;
    if (sc->physics == GD2_MHD) return mhd_wavespeed(g, U, wmax_bits, st);
    hipLaunchKernelGGL(euler_wavespeed_kernel, dim3(ew_blocks(n)), dim3(256), 0, st, U, n, n, g->gamma,
                       g->gamma - 1.0, 1.0 / g->dx, 1.0 / g->dy, wmax_bits);
    GD2_HIP(hipGetLastError());
    return GD2_OK;
}

static int ref_euler(const gd2_grid* g, int order, const gd2_double2* s, const gd2_double2* u, gd2_double2* s2,
                     gd2_double2* u2, double* T, const double* T_do, int method, void* stream) {
    // Kernel_TVD_old (order 2) takes the sphere, the 1/y terms and its west
    // ghost, Ker_Dekard none
    if (!grid_ok(g, true, order == 2 ? kProdOptions | kProdWestInflow : 0) || !s || !u || !s2 || !u2 || !T ||
        !T_do)
        return GD2_ERR_ARG;
    if ((const void*)s == (const void*)s2 || (const void*)u == (const void*)u2) return GD2_ERR_ARG;
    if (method < 0 || method > 4) return GD2_ERR_UNSUPPORTED;
    EulerArgs A;
    fill_common(A, g);
    A.s = reinterpret_cast<const double2*>(s);
    A.u = reinterpret_cast<const double2*>(u);
    A.s2 = reinterpret_cast<double2*>(s2);
    A.u2 = reinterpret_cast<double2*>(u2);
    A.dt_dev = T_do;
    A.tmin_bits = reinterpret_cast<unsigned long long*>(T);
    A.a = 0.0;
    A.b = 1.0;
    A.x0 = g->x0;
    A.y0 = g->y0;
    A.r_frozen = g->r_frozen;
    A.axisym = g->axisym ? 1 : 0;
    const bool west_in = g->bc[0] == GD2_BC_WEST_INFLOW;
    A.p_inflow = west_in ? 1.0 / (g->gamma * g->m_inf * g->m_inf) : 0.0;  // kernel.cu:3326
    if (!A.diag) return GD2_ERR_HIP;
    // methods 3 and 4: the HLLC variants HLLC_Korolkov and HLLC_Korolkov_2D
    // (no reference kernel calls them)
    static const int method_flux[5] = {GD2_FLUX_HLL, GD2_FLUX_HLLC, GD2_FLUX_EXACT, GD2_FLUX_HLLC_KOROLKOV,
                                       GD2_FLUX_HLLC_2D};
    const int flux = method_flux[method];
    if (order == 2 && west_in) {
        const hipStream_t st = (hipStream_t)stream;
        if (is_hllc_variant(flux)) return launch_euler_variant(2, flux, kAoS, false, false, kRefWestInflow, A, st);
        if (flux == GD2_FLUX_HLLC) return launch_euler<2, GD2_FLUX_HLLC, kAoS, false, false, kRefWestInflow>(A, st);
        if (flux == GD2_FLUX_EXACT) return launch_euler<2, GD2_FLUX_EXACT, kAoS, false, false, kRefWestInflow>(A, st);
        return launch_euler<2, GD2_FLUX_HLL, kAoS, false, false, kRefWestInflow>(A, st);
    }
    if (order == 2 && (A.axisym || A.r_frozen > 0.0)) {
        const hipStream_t st = (hipStream_t)stream;
        if (is_hllc_variant(flux)) return launch_euler_variant(2, flux, kAoS, false, false, kRefOptions, A, st);
        if (flux == GD2_FLUX_HLLC) return launch_euler<2, GD2_FLUX_HLLC, kAoS, false, false, kRefOptions>(A, st);
        if (flux == GD2_FLUX_EXACT) return launch_euler<2, GD2_FLUX_EXACT, kAoS, false, false, kRefOptions>(A, st);
        return launch_euler<2, GD2_FLUX_HLL, kAoS, false, false, kRefOptions>(A, st);
    }
    return dispatch_euler<kAoS, false, false>(order, flux, A, (hipStream_t)stream);
}

int gd2_ker_dekard(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, gd2_double2* s2, gd2_double2* u2,
                   double* T, const double* T_do, int method, void* stream) {
    return ref_euler(g, 1, s, u, s2, u2, T, T_do, method, stream);
}

int gd2_kernel_tvd_old(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, gd2_double2* s2,
                       gd2_double2* u2, double* T, const double* T_do, int method, void* stream) {
    return ref_euler(g, 2, s, u, s2, u2, T, T_do, method, stream);
}

int gd2_add2(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b, gd2_double2* s2,
             gd2_double2* u2, gd2_double3* b2, double* T, const double* T_do, int i, int method, void* stream) {
    (void)i;
    if (!grid_ok(g, true, kProdAll)) return GD2_ERR_ARG;
    return mhd_ref_step(g, 1, s, u, b, s2, u2, b2, T, T_do, method, diag_ptr(), (hipStream_t)stream);
}

int gd2_kernel_tvd(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b,
                   gd2_double2* s2, gd2_double2* u2, gd2_double3* b2, double* T, const double* T_do, int i,
                   int method, void* stream) {
    (void)i;
    if (!grid_ok(g, true, kProdAll)) return GD2_ERR_ARG;
    return mhd_ref_step(g, 2, s, u, b, s2, u2, b2, T, T_do, method, diag_ptr(), (hipStream_t)stream);
}

int gd2_funk_time(double* T, double* T_do, double* TT, int* i, void* stream) {
    if (!T || !T_do || !TT || !i) return GD2_ERR_ARG;
    hipLaunchKernelGGL(funk_time_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, T, T_do, TT, i);
    GD2_HIP(hipGetLastError());
    return GD2_OK;
}

int gd2_prim_to_cons(const gd2_grid* g, int physics, const gd2_double2* s, const gd2_double2* u,
                     const gd2_double3* b, double* U, void* stream) {
    if (!grid_ok(g, true, kProdAll) || !s || !u || !U || (physics == GD2_MHD && !b)) return GD2_ERR_ARG;
    if (physics != GD2_EULER && physics != GD2_MHD) return GD2_ERR_ARG;
    const long long n = (long long)g->nx * g->ny;
    hipLaunchKernelGGL(prim_to_cons_kernel, dim3(ew_blocks(n)), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const double2*>(s), reinterpret_cast<const double2*>(u),
                       reinterpret_cast<const double3*>(b), U, n, g->gamma, physics == GD2_MHD ? 1 : 0);
    GD2_HIP(hipGetLastError());
    return GD2_OK;
}

int gd2_cons_to_prim(const gd2_grid* g, int physics, const double* U, gd2_double2* s, gd2_double2* u,
                     gd2_double3* b, void* stream) {
    if (!grid_ok(g, true, kProdAll) || !s || !u || !U || (physics == GD2_MHD && !b)) return GD2_ERR_ARG;
    if (physics != GD2_EULER && physics != GD2_MHD) return GD2_ERR_ARG;
    const long long n = (long long)g->nx * g->ny;
    hipLaunchKernelGGL(cons_to_prim_kernel, dim3(ew_blocks(n)), dim3(256), 0, (hipStream_t)stream, U,
                       reinterpret_cast<double2*>(s), reinterpret_cast<double2*>(u), reinterpret_cast<double3*>(b), n,
                       g->gamma, physics == GD2_MHD ? 1 : 0);
    GD2_HIP(hipGetLastError());
    return GD2_OK;
}

int gd2_ctx_create(const gd2_grid* g, const gd2_scheme* sc, gd2_ctx** out) {
    if (!out) return GD2_ERR_ARG;
    *out = nullptr;
    if (!grid_ok(g, false) || !scheme_ok(sc)) return GD2_ERR_ARG;
    if (sc->integrator != GD2_FORWARD_EULER && sc->integrator != GD2_SSPRK2) return GD2_ERR_ARG;
    if (!(sc->cfl > 0.0)) return GD2_ERR_ARG;
    if (!diag_ptr()) return GD2_ERR_HIP;
    gd2_ctx* c = new (std::nothrow) gd2_ctx();
    if (!c) return GD2_ERR_NOMEM;
    c->g = *g;
    c->sc = *sc;
    const size_t bytes = sizeof(double) * (size_t)gd2_nfields(sc->physics) * (size_t)g->nx * (size_t)g->ny;
    hipError_t e = hipMalloc(&c->scratch, bytes);
    if (e != hipSuccess) {
        delete c;
        g_last_hip = (int)e;
        return GD2_ERR_NOMEM;
    }
    e = hipMalloc(&c->clock, sizeof(Clock));
    if (e != hipSuccess) {
        (void)hipFree(c->scratch);
        delete c;
        g_last_hip = (int)e;
        return GD2_ERR_NOMEM;
    }
    c->host_steps = 0;
    c->dt_log = nullptr;
    c->dt_cap = 0;
    *out = c;
    return GD2_OK;
}

int gd2_ctx_time_kernels(gd2_ctx* c, int capacity, int stride) {
    if (!c || capacity < 0 || stride < 1) return GD2_ERR_ARG;
    drop_events(c);
    if (capacity == 0) return GD2_OK;
    c->ev = new hipEvent_t[2 * (size_t)capacity];
    for (int i = 0; i < 2 * capacity; ++i) {
        const hipError_t e = hipEventCreate(&c->ev[i]);
        if (e != hipSuccess) {
            for (int k = 0; k < i; ++k) (void)hipEventDestroy(c->ev[k]);
            delete[] c->ev;
            c->ev = nullptr;
            g_last_hip = (int)e;
            return GD2_ERR_HIP;
        }
    }
    c->ev_cap = capacity;
    c->ev_n = 0;
    c->ev_stride = stride;
    c->ev_seen = 0;
    return GD2_OK;
}

int gd2_ctx_kernel_ms(gd2_ctx* c, double* mean_ms, int* count) {
    if (!c || !mean_ms) return GD2_ERR_ARG;
    double sum = 0.0;
    for (int i = 0; i < c->ev_n; ++i) {
        float ms = 0.0f;
        GD2_HIP(hipEventSynchronize(c->ev[2 * i + 1]));
        GD2_HIP(hipEventElapsedTime(&ms, c->ev[2 * i], c->ev[2 * i + 1]));
        sum += ms;
    }
    *mean_ms = c->ev_n > 0 ? sum / c->ev_n : 0.0;
    if (count) *count = c->ev_n;
    c->ev_n = 0;
    return GD2_OK;
}

int gd2_ctx_log_dt(gd2_ctx* c, double* dt_log, long long capacity) {
    if (!c || capacity < 0 || (capacity > 0 && !dt_log)) return GD2_ERR_ARG;
    c->dt_log = capacity > 0 ? dt_log : nullptr;
    c->dt_cap = capacity;
    return GD2_OK;
}

long long gd2_vjp_workspace_bytes(const gd2_grid* g, const gd2_scheme* sc) {
    if (!g || !sc || g->nx < 1 || g->ny < 1) return GD2_ERR_ARG;
    if (sc->physics == GD2_MHD) {
        // the cells' x-gathers (7 planes), the y-face-state cotangents (14
        // planes over the y-faces), the gathers of the 4 ghost columns and the
        // 4 ghost rows (7 each)
        const long long cells = (long long)g->nx * g->ny, nfy = (long long)g->nx * (g->ny + 1);
        return (long long)sizeof(double) * (7 * cells + 14 * nfy + 28LL * (g->nx + g->ny));
    }
    const long long len = g->nx > g->ny ? g->nx : g->ny;
    return (long long)sizeof(double) * 4 * 2 * len * 4;
}

int gd2_stage_vjp(const gd2_grid* g, const gd2_scheme* sc, const double* U, const double* V, const double* A,
                  double a, const double* B, double b, double c, double* W, void* workspace, void* stream) {
    if (!grid_ok(g, false) || !scheme_ok(sc) || !U || !V || !W || !workspace) return GD2_ERR_ARG;
    if (W == U || W == V) return GD2_ERR_ARG;
    if (sc->physics == GD2_MHD)
        return mhd_stage_vjp(g, sc, U, V, A, a, B, b, c, W, workspace, (hipStream_t)stream);
    if (sc->physics != GD2_EULER || sc->flux != GD2_FLUX_HLLC) return GD2_ERR_UNSUPPORTED;
    AdjArgs J;
    std::memset(&J, 0, sizeof J);
    J.U = U;
    J.V = V;
    J.A = A;
    J.B = B;
    J.W = W;
    J.a = a;
    J.b = b;
    J.c = c;
    J.ghost = static_cast<double*>(workspace);
    J.nx = g->nx;
    J.ny = g->ny;
    J.plane = (long long)g->nx * g->ny;
    J.dx = g->dx;
    J.dy = g->dy;
    J.idx = 1.0 / g->dx;
    J.idy = 1.0 / g->dy;
    J.gamma = g->gamma;
    J.g1 = g->gamma - 1.0;
    J.inv_g1 = 1.0 / (g->gamma - 1.0);
    for (int s2 = 0; s2 < 4; ++s2) J.bc[s2] = g->bc[s2];
    J.tiles_x = (g->nx + kAdjW - 1) / kAdjW;
    J.rows = strip_rows(sc->order == 2 ? reinterpret_cast<const void*>(&euler_vjp_kernel<2>)
                                       : reinterpret_cast<const void*>(&euler_vjp_kernel<1>),
                        J.tiles_x, g->ny, 3);
    {
        // GD2_VJP_ROWS overrides the strip height of the VJP launches alone
        static int env = -1;
        if (env < 0) {
            const char* e = std::getenv("GD2_VJP_ROWS");
            env = e ? std::atoi(e) : 0;
        }
        if (env > 0) J.rows = env;
    }
    J.gstride = g->nx > g->ny ? g->nx : g->ny;
    // ghost cotangents are folded inside the kernel when every ghost copies a
    // cell of its own wave or strip: no periodic side, and the first / last
    // column tile and row strip hold at least two columns / rows
    {
        bool in_kernel = g->nx >= 2 && g->ny >= 2 && J.rows >= 2;
        for (int s2 = 0; s2 < 4; ++s2) in_kernel = in_kernel && g->bc[s2] != GD2_BC_PERIODIC;
        in_kernel = in_kernel && g->nx - (J.tiles_x - 1) * kAdjW >= 2;
        const int nstrips = (g->ny + J.rows - 1) / J.rows;
        in_kernel = in_kernel && g->ny - (nstrips - 1) * J.rows >= 2;
        J.fold_in_kernel = in_kernel ? 1 : 0;
    }
    hipStream_t st = (hipStream_t)stream;
    const unsigned blocks = (unsigned)(J.tiles_x * ((g->ny + J.rows - 1) / J.rows));
    if (sc->order == 2)
        hipLaunchKernelGGL(euler_vjp_kernel<2>, dim3(blocks), dim3(kWave), 0, st, J);
    else
        hipLaunchKernelGGL(euler_vjp_kernel<1>, dim3(blocks), dim3(kWave), 0, st, J);
    GD2_HIP(hipGetLastError());
    if (!J.fold_in_kernel) {
        const long long nfold = 4ll * g->nx + 4ll * g->ny;
        hipLaunchKernelGGL(adj_fold_kernel, dim3((unsigned)((nfold + 255) / 256)), dim3(256), 0, st, J);
        GD2_HIP(hipGetLastError());
    }
    return GD2_OK;
}

int gd2_stage_vjp_noalloc(const gd2_grid* g, const gd2_scheme* sc, const double* U, const double* V,
                          const double* A, double a, const double* B, double b, double c, double* W,
                          void* stream) {
    // the same checks as gd2_stage_vjp, then the size limit, all before any HIP call
    if (!grid_ok(g, false) || !scheme_ok(sc) || !U || !V || !W) return GD2_ERR_ARG;
    if (W == U || W == V) return GD2_ERR_ARG;
    if (sc->physics == GD2_EULER && sc->flux != GD2_FLUX_HLLC) return GD2_ERR_UNSUPPORTED;
    const long long need = gd2_vjp_workspace_bytes(g, sc);
    if (need < 0) return GD2_ERR_ARG;
    if (need > GD2_NOALLOC_SCRATCH_BYTES) return GD2_ERR_TOO_LARGE;
    // every workspace entry the kernels read is written earlier in the same
    // call (ghost cotangents, face-state cotangents), so the contents left by
    // the last call need no clearing
    void* ws = noalloc_scratch();
    if (!ws) return GD2_ERR_HIP;
    return gd2_stage_vjp(g, sc, U, V, A, a, B, b, c, W, ws, stream);
}

int gd2_ctx_destroy(gd2_ctx* c) {
    if (!c) return GD2_ERR_ARG;
    (void)hipDeviceSynchronize();
    drop_events(c);
    (void)hipFree(c->scratch);
    (void)hipFree(c->clock);
    delete c;
    return GD2_OK;
}

int gd2_ctx_start(gd2_ctx* c, const double* U, double t0, void* stream) {
    if (!c || !U) return GD2_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(clock_start_kernel, dim3(1), dim3(1), 0, st, c->clock, t0, c->sc.cfl);
    GD2_HIP(hipGetLastError());
    c->host_steps = 0;
    return gd2_wavespeed_max(&c->g, &c->sc, U, &c->clock->w[0], stream);
}

// How SSP-RK2 Euler steps are taken: GD2_STEP_AUTO picks the one-pass kernel
// on grids of at least kOnePassCells cells, where it is measured faster than
// two stage launches (on MI355X: 0.80x at 4096^2, 0.90x at 2048^2, 1.05x at
// 1024^2, 1.17x at 512^2: a small grid needs short strips to fill the chip,
// and the one-pass strip re-forms four more rows of U1 than a stage strip
// re-reads). Both paths give the same bits. GD2_TWO_PASS=1 in the
// environment starts in GD2_STEP_TWO_PASS.
constexpr long long kOnePassCells = 1ll << 21;
static std::atomic<int> g_step_path{-1};

static int step_path() {
    int v = g_step_path.load();
    if (v < 0) {
        const char* e = std::getenv("GD2_TWO_PASS");
        int init = (e && std::atoi(e) == 1) ? GD2_STEP_TWO_PASS : GD2_STEP_AUTO;
        g_step_path.compare_exchange_strong(v, init);
        v = g_step_path.load();
    }
    return v;
}

int gd2_set_step_path(int mode) {
    if (mode != GD2_STEP_AUTO && mode != GD2_STEP_TWO_PASS && mode != GD2_STEP_ONE_PASS) return GD2_ERR_ARG;
    const int prev = step_path();
    g_step_path.store(mode);
    return prev;
}

int gd2_one_pass_preferred(const gd2_grid* g, const gd2_scheme* sc) {
    if (!g || !sc) return GD2_ERR_ARG;
    if (sc->physics != GD2_EULER || sc->integrator != GD2_SSPRK2 || g->nx < 2 || g->ny < 2) return 0;
    const int mode = step_path();
    if (mode == GD2_STEP_ONE_PASS) return 1;
    if (mode == GD2_STEP_TWO_PASS) return 0;
    return (long long)g->nx * g->ny >= kOnePassCells ? 1 : 0;
}

int gd2_ctx_advance(gd2_ctx* c, double* U, int nsteps, double t_end, void* stream) {
    if (!c || !U || nsteps < 0) return GD2_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    double* S = c->scratch;
    const double* dtp = &c->clock->dt;
    double* cur = U;  // forward Euler and the one-pass SSP-RK2 step ping-pong between U and S
    const bool one_pass = gd2_one_pass_preferred(&c->g, &c->sc) == 1;
    // events on a stream under capture would become capture dependencies,
    // not timestamps: no kernel timing while the stream is captured
    bool capturing = false;
    if (c->ev_cap > 0) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GD2_HIP(hipStreamIsCapturing(st, &cs));
        capturing = cs != hipStreamCaptureStatusNone;
    }
    // the one-pass step with no end time does its own clock update (one
    // launch per step); otherwise a one-thread clock kernel starts the step
    const bool fused_clock = c->sc.integrator == GD2_SSPRK2 && one_pass && !(t_end > 0.0);
    for (int i = 0; i < nsteps; ++i) {
        const int slot = (int)(c->host_steps % 3);
        if (!fused_clock) {
            hipLaunchKernelGGL(clock_begin_kernel, dim3(1), dim3(1), 0, st, c->clock, slot, t_end, c->dt_log,
                               c->dt_cap);
            GD2_HIP(hipGetLastError());
        }
        unsigned long long* wn = &c->clock->w[slot == 2 ? 0 : slot + 1];
        const bool timed = !capturing && c->ev_n < c->ev_cap && c->ev_seen++ % c->ev_stride == 0;
        if (timed) GD2_HIP(hipEventRecord(c->ev[2 * c->ev_n], st));
        int rc;
        if (c->sc.integrator == GD2_SSPRK2 && one_pass) {
            // one pass per step, ping-ponging between U and the scratch state
            double* nxt = (cur == U) ? S : U;
            if (fused_clock)
                rc = ssprk2_launch(&c->g, &c->sc, cur, nxt, nullptr, 0.0, nullptr, wn, c->clock, slot, c->dt_log,
                                   c->dt_cap, st);
            else
                rc = gd2_ssprk2_step(&c->g, &c->sc, cur, nxt, nullptr, 0.0, dtp, wn, stream);
            if (rc) return rc;
            cur = nxt;
        } else if (c->sc.integrator == GD2_SSPRK2) {
            rc = gd2_stage(&c->g, &c->sc, U, nullptr, S, 0.0, 1.0, 0.0, dtp, nullptr, stream);
            if (rc) return rc;
            rc = gd2_stage(&c->g, &c->sc, S, U, U, 0.5, 0.5, 0.0, dtp, wn, stream);
            if (rc) return rc;
        } else {
            double* nxt = (cur == U) ? S : U;
            rc = gd2_stage(&c->g, &c->sc, cur, nullptr, nxt, 0.0, 1.0, 0.0, dtp, wn, stream);
            if (rc) return rc;
            cur = nxt;
        }
        if (timed) {
            GD2_HIP(hipEventRecord(c->ev[2 * c->ev_n + 1], st));
            c->ev_n += 1;
        }
        c->host_steps += 1;
    }
    if (cur != U) {
        const size_t bytes = sizeof(double) * (size_t)gd2_nfields(c->sc.physics) * (size_t)c->g.nx * (size_t)c->g.ny;
        GD2_HIP(hipMemcpyAsync(U, cur, bytes, hipMemcpyDeviceToDevice, st));
    }
    return GD2_OK;
}

int gd2_ctx_time(gd2_ctx* c, double* t, long long* steps, double* dt, void* stream) {
    if (!c) return GD2_ERR_ARG;
    Clock h;
    GD2_HIP(hipMemcpyAsync(&h, c->clock, sizeof(Clock), hipMemcpyDeviceToHost, (hipStream_t)stream));
    GD2_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (t) *t = h.t;
    if (steps) *steps = h.step;
    if (dt) *dt = h.dt;
    return GD2_OK;
}

int gd2_get_diagnostics(gd2_diag* out, int reset) {
    DeviceDiag* p = diag_ptr();
    if (!p || !out) return p ? GD2_ERR_ARG : GD2_ERR_HIP;
    GD2_HIP(hipDeviceSynchronize());
    DeviceDiag h;
    GD2_HIP(hipMemcpy(&h, p, sizeof h, hipMemcpyDeviceToHost));
    out->clamp_rho = (long long)h.clamp_rho;
    out->clamp_p = (long long)h.clamp_p;
    out->newton_cap_hits = (long long)h.newton_cap;
    out->bisect_cap_hits = (long long)h.bisect_cap;
    if (reset) {
        std::memset(&h, 0, sizeof h);
        GD2_HIP(hipMemcpy(p, &h, sizeof h, hipMemcpyHostToDevice));
    }
    return GD2_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- ensembles

struct gd2_ensemble {
    int n = 0;
    gd2_scheme sc{};
    std::vector<EnsProblem> host;     // descriptors (device pointers inside)
    EnsProblem* dprobs = nullptr;
    Clock* clocks = nullptr;
    double* scratch = nullptr;
    EnsTile* dtiles = nullptr;        // stage tiles (an odd step count's first step)
    int ntiles = 0;
    EnsTile* dtiles1 = nullptr;       // one-pass tiles (the persistent launch), by work list
    int ntiles1 = 0;
    EnsQueues q_xcd{}, q_one{};       // XCD-affine work lists / one list
    bool xcd_ok = false;              // the XCD bins balance
    unsigned* sync = nullptr;         // work-list heads, spin flag, per-problem completions
    unsigned long long* hist = nullptr;  // per-problem wave-speed maxima of a launch's steps
    int rows = 0;       // one-pass strip height asked for (0 = default)
    int rows1 = 0;      // one-pass strip height in use
    int blocks = 0;     // waves of the persistent launch (0 = all that fit)
    int resident = 0;   // waves of the persistent kernel that fit on the device
    long long host_steps = 0;
    hipStream_t cap_stream = nullptr;
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;  // (nsteps, slot) -> graph
    std::vector<hipGraph_t> raw_graphs;
};

static void ens_drop_graphs(gd2_ensemble* e) {
    for (auto& kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto g : e->raw_graphs) (void)hipGraphDestroy(g);
    e->graphs.clear();
    e->raw_graphs.clear();
}

// one-pass strip height: the requested one, else the tallest of 256, 128, ...
// 16 that still gives the launch two tiles per resident wave
static int ens_rk2_rows(const gd2_ensemble* e) {
    if (e->rows > 0) return std::max(2, e->rows);  // strips start at row 0 or at row >= 2
    int rows = 256;
    for (; rows > 16; rows /= 2) {
        long long t = 0;
        for (int p = 0; p < e->n; ++p) {
            const EulerArgs& A = e->host[p].args;
            t += (long long)((A.nx + kRk2W - 1) / kRk2W) * ((A.ny + rows - 1) / rows);
        }
        if (t >= 2LL * std::max(e->resident, 1)) break;
    }
    return rows;
}

// large problems first, so the end of each step's share of the list holds
// small tiles; the tiles of one problem stay together
static void ens_sort(const gd2_ensemble* e, std::vector<EnsTile>& t) {
    std::stable_sort(t.begin(), t.end(), [&](const EnsTile& a, const EnsTile& b) {
        const long long ca = (long long)e->host[a.problem].args.nx * e->host[a.problem].args.ny;
        const long long cb = (long long)e->host[b.problem].args.nx * e->host[b.problem].args.ny;
        return ca > cb;
    });
}

static int ens_build_tiles(gd2_ensemble* e) {
    std::vector<EnsTile> t;
    const int rows1 = ens_rk2_rows(e);
    for (int p = 0; p < e->n; ++p) {
        EnsProblem& P = e->host[p];
        EulerArgs& A = P.args;
        A.rows = tile_rows_for(A.nx, A.ny);
        const int ty_n = (A.ny + A.rows - 1) / A.rows;
        for (int ty = 0; ty < ty_n; ++ty)
            for (int tx = 0; tx < A.tiles_x; ++tx) t.push_back(EnsTile{p, tx, ty, 0});
        P.rk2_rows = rows1;
        P.rk2_tiles = ((A.nx + kRk2W - 1) / kRk2W) * ((A.ny + rows1 - 1) / rows1);
    }
    ens_sort(e, t);
    // XCD bins: largest work first into the least loaded bin; a problem's work
    // is its strips' passes (rows plus about five) times its strip columns
    std::vector<long long> cost(e->n);
    std::vector<int> order(e->n), bin(e->n);
    for (int p = 0; p < e->n; ++p) {
        const EulerArgs& A = e->host[p].args;
        long long passes = 0;
        for (int j0 = 0; j0 < A.ny; j0 += rows1) passes += std::min(rows1, A.ny - j0) + 5;
        cost[p] = passes * ((A.nx + kRk2W - 1) / kRk2W);
        order[p] = p;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    long long load[kEnsQueues] = {0}, total = 0;
    for (int p : order) {
        int b = 0;
        for (int k = 1; k < kEnsQueues; ++k)
            if (load[k] < load[b]) b = k;
        bin[p] = b;
        load[b] += cost[p];
        total += cost[p];
    }
    long long worst = 0;
    for (int k = 0; k < kEnsQueues; ++k) worst = std::max(worst, load[k]);
    e->xcd_ok = worst * kEnsQueues <= total + total / 16;  // within about 6 % of an even split
    // one-pass tiles by bin, then large problems first (the tiles of one
    // problem stay together)
    std::vector<EnsTile> t1;
    for (int k = 0; k < kEnsQueues; ++k) {
        e->q_xcd.off[k] = (int)t1.size();
        std::vector<EnsTile> tb;
        for (int p = 0; p < e->n; ++p) {
            if (bin[p] != k) continue;
            const EulerArgs& A = e->host[p].args;
            const int tx1 = (A.nx + kRk2W - 1) / kRk2W, ty1 = (A.ny + rows1 - 1) / rows1;
            for (int ty = 0; ty < ty1; ++ty)
                for (int tx = 0; tx < tx1; ++tx) tb.push_back(EnsTile{p, tx, ty, 0});
        }
        ens_sort(e, tb);
        t1.insert(t1.end(), tb.begin(), tb.end());
    }
    e->q_xcd.off[kEnsQueues] = (int)t1.size();
    e->q_one.off[0] = 0;
    for (int k = 1; k <= kEnsQueues; ++k) e->q_one.off[k] = (int)t1.size();
    if (e->dtiles) (void)hipFree(e->dtiles);
    if (e->dtiles1) (void)hipFree(e->dtiles1);
    e->dtiles = e->dtiles1 = nullptr;
    GD2_HIP(hipMalloc(&e->dtiles, sizeof(EnsTile) * t.size()));
    GD2_HIP(hipMalloc(&e->dtiles1, sizeof(EnsTile) * t1.size()));
    GD2_HIP(hipMemcpy(e->dtiles, t.data(), sizeof(EnsTile) * t.size(), hipMemcpyHostToDevice));
    GD2_HIP(hipMemcpy(e->dtiles1, t1.data(), sizeof(EnsTile) * t1.size(), hipMemcpyHostToDevice));
    GD2_HIP(hipMemcpy(e->dprobs, e->host.data(), sizeof(EnsProblem) * e->n, hipMemcpyHostToDevice));
    e->ntiles = (int)t.size();
    e->ntiles1 = (int)t1.size();
    e->rows1 = rows1;
    ens_drop_graphs(e);
    return GD2_OK;
}

template <int STAGE>
static int ens_launch_stage(gd2_ensemble* e, int slot, hipStream_t st) {
    const unsigned grid = (unsigned)e->ntiles;
    const int order = e->sc.order, flux = e->sc.flux;
    if (order == 2 && flux == GD2_FLUX_HLLC)
        hipLaunchKernelGGL((ens_stage_kernel<2, GD2_FLUX_HLLC, STAGE>), dim3(grid), dim3(kWave), 0, st, e->dprobs,
                           e->dtiles, e->ntiles, slot);
    else if (order == 1 && flux == GD2_FLUX_HLLC)
        hipLaunchKernelGGL((ens_stage_kernel<1, GD2_FLUX_HLLC, STAGE>), dim3(grid), dim3(kWave), 0, st, e->dprobs,
                           e->dtiles, e->ntiles, slot);
    else if (order == 2 && flux == GD2_FLUX_EXACT)
        hipLaunchKernelGGL((ens_stage_kernel<2, GD2_FLUX_EXACT, STAGE>), dim3(grid), dim3(kWave), 0, st, e->dprobs,
                           e->dtiles, e->ntiles, slot);
    else if (order == 1 && flux == GD2_FLUX_EXACT)
        hipLaunchKernelGGL((ens_stage_kernel<1, GD2_FLUX_EXACT, STAGE>), dim3(grid), dim3(kWave), 0, st, e->dprobs,
                           e->dtiles, e->ntiles, slot);
    else if (order == 2)
        hipLaunchKernelGGL((ens_stage_kernel<2, GD2_FLUX_HLL, STAGE>), dim3(grid), dim3(kWave), 0, st, e->dprobs,
                           e->dtiles, e->ntiles, slot);
    else
        hipLaunchKernelGGL((ens_stage_kernel<1, GD2_FLUX_HLL, STAGE>), dim3(grid), dim3(kWave), 0, st, e->dprobs,
                           e->dtiles, e->ntiles, slot);
    GD2_HIP(hipGetLastError());
    return GD2_OK;
}

static const void* ens_rk2_symbol(int order, int flux) {
    if (order == 2 && flux == GD2_FLUX_HLLC) return reinterpret_cast<const void*>(&ens_rk2_kernel<2, GD2_FLUX_HLLC, true>);
    if (order == 1 && flux == GD2_FLUX_HLLC) return reinterpret_cast<const void*>(&ens_rk2_kernel<1, GD2_FLUX_HLLC, true>);
    if (order == 2 && flux == GD2_FLUX_EXACT)
        return reinterpret_cast<const void*>(&ens_rk2_kernel<2, GD2_FLUX_EXACT, true>);
    if (order == 1 && flux == GD2_FLUX_EXACT)
        return reinterpret_cast<const void*>(&ens_rk2_kernel<1, GD2_FLUX_EXACT, true>);
    if (order == 2) return reinterpret_cast<const void*>(&ens_rk2_kernel<2, GD2_FLUX_HLL, true>);
    return reinterpret_cast<const void*>(&ens_rk2_kernel<1, GD2_FLUX_HLL, true>);
}

template <bool XCD>
static void ens_rk2_go(gd2_ensemble* e, unsigned grid, const EnsQueues& q, int nsteps, int dry, hipStream_t st) {
    const int order = e->sc.order, flux = e->sc.flux;
#define GD2_ENS_GO(O, F)                                                                                           \
    hipLaunchKernelGGL((ens_rk2_kernel<O, F, XCD>), dim3(grid), dim3(kWave), 0, st, e->dprobs, e->dtiles1, q, nsteps, \
                       e->hist, e->sync, dry)
    if (order == 2 && flux == GD2_FLUX_HLLC)
        GD2_ENS_GO(2, GD2_FLUX_HLLC);
    else if (order == 1 && flux == GD2_FLUX_HLLC)
        GD2_ENS_GO(1, GD2_FLUX_HLLC);
    else if (order == 2 && flux == GD2_FLUX_EXACT)
        GD2_ENS_GO(2, GD2_FLUX_EXACT);
    else if (order == 1 && flux == GD2_FLUX_EXACT)
        GD2_ENS_GO(1, GD2_FLUX_EXACT);
    else if (order == 2)
        GD2_ENS_GO(2, GD2_FLUX_HLL);
    else
        GD2_ENS_GO(1, GD2_FLUX_HLL);
#undef GD2_ENS_GO
}

// persistent launches of nsteps one-pass steps from clock slot `slot`, up to
// kEnsMaxSteps per launch, each between a reset and (unless dry) a clock
// finish kernel
static int ens_launch_rk2(gd2_ensemble* e, int nsteps, int slot, int dry, hipStream_t st) {
    const int want = e->blocks > 0 ? e->blocks : e->resident;
    // XCD-affine lists need waves on every XCD: launches of at least 64 waves
    // (consecutive workgroups go to the XCDs in turn)
    const bool xcd = e->xcd_ok && want >= 8 * kEnsQueues;
    const int nr = kEnsSyncHead + e->n;
    while (nsteps > 0) {
        const int k = std::min(nsteps, kEnsMaxSteps);
        hipLaunchKernelGGL(ens_launch_reset_kernel, dim3((nr + 255) / 256), dim3(256), 0, st, e->dprobs, e->n, k, slot,
                           e->hist, e->sync);
        GD2_HIP(hipGetLastError());
        const long long items = (long long)e->ntiles1 * k;
        const unsigned grid = (unsigned)std::max(1LL, std::min<long long>(want, xcd ? std::max<long long>(items, 64) : items));
        if (xcd)
            ens_rk2_go<true>(e, grid, e->q_xcd, k, dry, st);
        else
            ens_rk2_go<false>(e, grid, e->q_one, k, dry, st);
        GD2_HIP(hipGetLastError());
        if (!dry) {
            hipLaunchKernelGGL(ens_launch_finish_kernel, dim3((e->n + 63) / 64), dim3(64), 0, st, e->dprobs, e->n, k,
                               slot, e->hist);
            GD2_HIP(hipGetLastError());
            slot = (slot + k) % 3;
        }
        nsteps -= k;
    }
    return GD2_OK;
}

// nsteps SSP-RK2 steps from clock slot first_slot: an odd count takes its
// first step as two stage launches (in place), the rest in one persistent
// launch that ping-pongs and ends in the caller's buffers
static int ens_enqueue_steps(gd2_ensemble* e, int nsteps, int first_slot, hipStream_t st) {
    int slot = first_slot;
    if (nsteps & 1) {
        hipLaunchKernelGGL(ens_clock_kernel, dim3((e->n + 63) / 64), dim3(64), 0, st, e->dprobs, e->n, slot);
        GD2_HIP(hipGetLastError());
        int rc = ens_launch_stage<1>(e, slot, st);
        if (rc) return rc;
        rc = ens_launch_stage<2>(e, slot, st);
        if (rc) return rc;
        slot = (slot + 1) % 3;
        nsteps -= 1;
    }
    if (nsteps > 0) return ens_launch_rk2(e, nsteps, slot, 0, st);
    return GD2_OK;
}

extern "C" {

int gd2_ensemble_create(int n, const gd2_grid* grids, const gd2_scheme* sc, double* const* U, gd2_ensemble** out) {
    if (!out) return GD2_ERR_ARG;
    *out = nullptr;
    if (n < 1 || !grids || !sc || !U || !scheme_ok(sc) || sc->physics != GD2_EULER) return GD2_ERR_ARG;
    // the ensemble's kernels are built for the HLLC, exact and HLL fluxes
    if (is_hllc_variant(sc->flux)) return GD2_ERR_UNSUPPORTED;
    if (sc->integrator != GD2_SSPRK2 || !(sc->cfl > 0.0)) return GD2_ERR_UNSUPPORTED;
    for (int p = 0; p < n; ++p) {
        if (!grid_ok(&grids[p], false) || !U[p]) return GD2_ERR_ARG;
        // the one-pass step builds the boundary rows and columns of its
        // intermediate state from the first and last two of them
        if (grids[p].nx < 2 || grids[p].ny < 2) return GD2_ERR_UNSUPPORTED;
    }
    if (!diag_ptr()) return GD2_ERR_HIP;
    gd2_ensemble* e = new (std::nothrow) gd2_ensemble();
    if (!e) return GD2_ERR_NOMEM;
    e->n = n;
    e->sc = *sc;
    e->host.resize(n);
    size_t total = 0;
    for (int p = 0; p < n; ++p) total += (size_t)4 * grids[p].nx * grids[p].ny;
    hipError_t err = hipMalloc(&e->scratch, sizeof(double) * total);
    if (err == hipSuccess) err = hipMalloc(&e->clocks, sizeof(Clock) * n);
    if (err == hipSuccess) err = hipMalloc(&e->dprobs, sizeof(EnsProblem) * n);
    // sync words: heads, flag, per-problem completions and per-problem totals
    if (err == hipSuccess) err = hipMalloc(&e->sync, sizeof(unsigned) * (kEnsSyncHead + 2 * n));
    if (err == hipSuccess) err = hipMemset(e->sync, 0, sizeof(unsigned) * (kEnsSyncHead + 2 * n));
    if (err == hipSuccess) err = hipMalloc(&e->hist, sizeof(unsigned long long) * kEnsHist * n);
    if (err != hipSuccess) {
        g_last_hip = (int)err;
        gd2_ensemble_destroy(e);
        return GD2_ERR_NOMEM;
    }
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ens_rk2_symbol(sc->order, sc->flux), kWave, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 8;
    e->resident = cus * per_cu;
    size_t off = 0;
    for (int p = 0; p < n; ++p) {
        EnsProblem& P = e->host[p];
        fill_common(P.args, &grids[p]);
        P.U = U[p];
        P.S = e->scratch + off;
        P.clk = e->clocks + p;
        P.clk_cfl = sc->cfl;
        off += (size_t)4 * grids[p].nx * grids[p].ny;
    }
    int rc = ens_build_tiles(e);
    if (rc) {
        gd2_ensemble_destroy(e);
        return rc;
    }
    *out = e;
    return GD2_OK;
}

int gd2_ensemble_destroy(gd2_ensemble* e) {
    if (!e) return GD2_ERR_ARG;
    (void)hipDeviceSynchronize();
    ens_drop_graphs(e);
    if (e->cap_stream) (void)hipStreamDestroy(e->cap_stream);
    if (e->dtiles) (void)hipFree(e->dtiles);
    if (e->dtiles1) (void)hipFree(e->dtiles1);
    if (e->sync) (void)hipFree(e->sync);
    if (e->hist) (void)hipFree(e->hist);
    if (e->dprobs) (void)hipFree(e->dprobs);
    if (e->clocks) (void)hipFree(e->clocks);
    if (e->scratch) (void)hipFree(e->scratch);
    delete e;
    return GD2_OK;
}

int gd2_ensemble_set_tiling(gd2_ensemble* e, int rows, int blocks) {
    if (!e || rows < 0 || blocks < 0 || rows > 4096) return GD2_ERR_ARG;
    e->rows = rows;
    e->blocks = blocks;
    return ens_build_tiles(e);
}

int gd2_ensemble_get_tiling(gd2_ensemble* e, int* rows, int* blocks, int* ntiles) {
    if (!e) return GD2_ERR_ARG;
    if (rows) *rows = e->rows1;
    if (blocks) *blocks = e->blocks > 0 ? e->blocks : e->resident;
    if (ntiles) *ntiles = e->ntiles1;
    return GD2_OK;
}

int gd2_ensemble_start(gd2_ensemble* e, void* stream) {
    if (!e) return GD2_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(ens_clock_start_kernel, dim3((e->n + 63) / 64), dim3(64), 0, st, e->clocks, e->n, e->sc.cfl);
    GD2_HIP(hipGetLastError());
    for (int p = 0; p < e->n; ++p) {
        const EulerArgs& A = e->host[p].args;
        const long long n = A.plane;
        hipLaunchKernelGGL(euler_wavespeed_kernel, dim3(ew_blocks(n)), dim3(256), 0, st, e->host[p].U, n, n, A.gamma,
                           A.g1, A.idx, A.idy, &e->clocks[p].w[0]);
        GD2_HIP(hipGetLastError());
    }
    e->host_steps = 0;
    return GD2_OK;
}

int gd2_ensemble_advance(gd2_ensemble* e, int nsteps, int use_graph, void* stream) {
    if (!e || nsteps < 0) return GD2_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (nsteps == 0) return GD2_OK;
    if (!use_graph) {
        int rc = ens_enqueue_steps(e, nsteps, (int)(e->host_steps % 3), st);
        if (rc) return rc;
        e->host_steps += nsteps;
        return GD2_OK;
    }
    const int parity = (int)(e->host_steps % 3);  // the first step's clock slot
    auto key = std::make_pair(nsteps, parity);
    auto it = e->graphs.find(key);
    if (it == e->graphs.end()) {
        if (!e->cap_stream) GD2_HIP(hipStreamCreateWithFlags(&e->cap_stream, hipStreamNonBlocking));
        GD2_HIP(hipStreamBeginCapture(e->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = ens_enqueue_steps(e, nsteps, parity, e->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(e->cap_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        GD2_HIP(ce);
        hipGraphExec_t ge = nullptr;
        GD2_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        e->raw_graphs.push_back(g);
        it = e->graphs.emplace(key, ge).first;
    }
    GD2_HIP(hipGraphLaunch(it->second, st));
    e->host_steps += nsteps;
    return GD2_OK;
}

int gd2_ensemble_time(gd2_ensemble* e, int p, double* t, long long* steps, double* dt, void* stream) {
    if (!e || p < 0 || p >= e->n) return GD2_ERR_ARG;
    Clock h;
    std::vector<unsigned> sy((size_t)kEnsSyncHead + 2 * e->n);
    GD2_HIP(hipMemcpyAsync(&h, e->clocks + p, sizeof(Clock), hipMemcpyDeviceToHost, (hipStream_t)stream));
    GD2_HIP(hipMemcpyAsync(sy.data(), e->sync, sizeof(unsigned) * sy.size(), hipMemcpyDeviceToHost,
                           (hipStream_t)stream));
    GD2_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (sy[kEnsFlag]) {  // a dependency wait reached its spin limit: report it once
        GD2_HIP(hipMemset(e->sync + kEnsFlag, 0, sizeof(unsigned)));
        return GD2_ERR_HIP;
    }
    for (int q = 0; q < e->n; ++q)  // the last persistent launch took every tile of every problem
        if (sy[kEnsSyncHead + q] != sy[kEnsSyncHead + e->n + q]) return GD2_ERR_HIP;
    if (t) *t = h.t;
    if (steps) *steps = h.step;
    if (dt) *dt = h.dt;
    return GD2_OK;
}

int gd2_ensemble_tune(gd2_ensemble* e, void* stream, int* rows_out, int* blocks_out, double* best_ms) {
    if (!e) return GD2_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    // candidate strip heights and persistent-launch sizes, each timed on a
    // dry launch of kTuneSteps steps (reads the states, writes only the
    // library's second buffers, leaves the clocks alone); launches of a few
    // steps favour short strips (their start and end weigh more), so the
    // dry launch is long enough to time the steady state
    constexpr int kTuneSteps = 16;
    const int rows_c[] = {32, 64, 128, 256};
    const int blocks_c[] = {(3 * e->resident) / 4, e->resident};
    const int slot = (int)(e->host_steps % 3);
    hipEvent_t a, b;
    GD2_HIP(hipEventCreate(&a));
    GD2_HIP(hipEventCreate(&b));
    double best = 1e300;
    int br = e->rows, bb = e->blocks;
    int rc = GD2_OK;
    for (int r : rows_c) {
        e->rows = r;
        rc = ens_build_tiles(e);
        if (rc) break;
        for (int blk : blocks_c) {
            if (blk < 1) continue;
            e->blocks = blk;
            rc = ens_launch_rk2(e, kTuneSteps, slot, 1, st);  // warm-up
            if (rc) break;
            GD2_HIP(hipEventRecord(a, st));
            rc = ens_launch_rk2(e, kTuneSteps, slot, 1, st);
            if (rc) break;
            GD2_HIP(hipEventRecord(b, st));
            GD2_HIP(hipEventSynchronize(b));
            float ms = 0.f;
            GD2_HIP(hipEventElapsedTime(&ms, a, b));
            if (ms < best) {
                best = ms;
                br = r;
                bb = blk;
            }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    e->rows = br;
    e->blocks = bb;
    const int rc2 = ens_build_tiles(e);
    if (rc) return rc;
    if (rc2) return rc2;
    if (rows_out) *rows_out = e->rows1;
    if (blocks_out) *blocks_out = e->blocks > 0 ? e->blocks : e->resident;
    if (best_ms) *best_ms = best / kTuneSteps;
    return GD2_OK;
}

}  // extern "C"

