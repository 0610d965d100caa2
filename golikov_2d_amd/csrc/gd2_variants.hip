// gd2_variants.hip -- the HLLC variants of the reference as selectable
// fluxes, and its interface fluxes evaluated pointwise on batches of faces.
//
// * Stage kernel and one-pass step instances for GD2_FLUX_HLLC_KOROLKOV
//   (HLLC_Korolkov, kernel.cu:1349) and GD2_FLUX_HLLC_2D (HLLC_Korolkov_2D,
//   Gas_Dinamic.cu:265), reached through gd2_stage, gd2_ssprk2_step, the
//   solver context and the reference-layout steps (methods 3 and 4).
// * gd2_euler_face_fluxes / gd2_mhd_face_fluxes: one GPU thread per face
//   evaluates one of the reference's flux functions (the device-function flux
//   interface of Header.h:41-65) with its own lab-frame arguments and outputs,
//   the passive-scalar variants included.
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/gd2.h"
#include "gd2_gkr.hpp"
#include "gd2_launch.hpp"
#include "gd2_mhd.hpp"

namespace gd2 {

template <int ORDER, int FLUX>
static int euler_variant(int layout, bool w, bool base, int prod, const EulerArgs& A, hipStream_t st) {
    if (layout == kAoS) {
        if (w || base) return GD2_ERR_UNSUPPORTED;
        if (prod != kRefPlain) {
            if constexpr (ORDER == 2) {
                if (prod == kRefWestInflow) return launch_euler<2, FLUX, kAoS, false, false, kRefWestInflow>(A, st);
                return launch_euler<2, FLUX, kAoS, false, false, kRefOptions>(A, st);
            }
            return GD2_ERR_UNSUPPORTED;
        }
        return launch_euler<ORDER, FLUX, kAoS, false, false>(A, st);
    }
    if (prod != kRefPlain) return GD2_ERR_UNSUPPORTED;
    if (w && base) return launch_euler<ORDER, FLUX, kSoA, true, true>(A, st);
    if (w) return launch_euler<ORDER, FLUX, kSoA, true, false>(A, st);
    if (base) return launch_euler<ORDER, FLUX, kSoA, false, true>(A, st);
    return launch_euler<ORDER, FLUX, kSoA, false, false>(A, st);
}

int launch_euler_variant(int order, int flux, int layout, bool w, bool base, int prod, const EulerArgs& A,
                         hipStream_t st) {
    if (flux == GD2_FLUX_HLLC_KOROLKOV) {
        if (order == 2) return euler_variant<2, GD2_FLUX_HLLC_KOROLKOV>(layout, w, base, prod, A, st);
        if (order == 1) return euler_variant<1, GD2_FLUX_HLLC_KOROLKOV>(layout, w, base, prod, A, st);
    } else if (flux == GD2_FLUX_HLLC_2D) {
        if (order == 2) return euler_variant<2, GD2_FLUX_HLLC_2D>(layout, w, base, prod, A, st);
        if (order == 1) return euler_variant<1, GD2_FLUX_HLLC_2D>(layout, w, base, prod, A, st);
    }
    return GD2_ERR_UNSUPPORTED;
}

int launch_rk2_variant(int order, int flux, bool keep, const EulerArgs& A, hipStream_t st) {
    if (flux == GD2_FLUX_HLLC_KOROLKOV) {
        if (order == 2)
            return keep ? launch_rk2<2, GD2_FLUX_HLLC_KOROLKOV, true>(A, st)
                        : launch_rk2<2, GD2_FLUX_HLLC_KOROLKOV, false>(A, st);
        if (order == 1)
            return keep ? launch_rk2<1, GD2_FLUX_HLLC_KOROLKOV, true>(A, st)
                        : launch_rk2<1, GD2_FLUX_HLLC_KOROLKOV, false>(A, st);
    } else if (flux == GD2_FLUX_HLLC_2D) {
        if (order == 2)
            return keep ? launch_rk2<2, GD2_FLUX_HLLC_2D, true>(A, st) : launch_rk2<2, GD2_FLUX_HLLC_2D, false>(A, st);
        if (order == 1)
            return keep ? launch_rk2<1, GD2_FLUX_HLLC_2D, true>(A, st) : launch_rk2<1, GD2_FLUX_HLLC_2D, false>(A, st);
    }
    return GD2_ERR_UNSUPPORTED;
}

// ------------------------------------------------------------ pointwise fluxes

struct FaceBatch {
    const double2 *Ls, *Lu, *Rs, *Ru;  // (rho, p), (vx, vy) of the left and right states
    const double *QL, *QR;             // passive scalar (Q kinds)
    const double2* nrm;                // face normal (n1, n2)
    double2 *Ps, *Pu;                  // (F_rho, F_E), (F_mx, F_my)
    double* PQ;                        // passive-scalar flux (Q kinds)
    double* dt;                        // face time step krit * rad / S
    long long n;
    double gamma, inv_g1, krit, rad;
    DeviceDiag* diag;
};

// One thread per face. The reference's functions project the velocities on
// n and t = (-n2, n1) and rotate the momentum flux back (kernel.cu:1187-1201,
// 1343-1344; Gas_Dinamic.cu:279-289); the device fluxes here work in that
// face frame.
template <int KIND>
__global__ __launch_bounds__(256) void euler_face_kernel(const FaceBatch B) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < B.n;
         i += (long long)gridDim.x * blockDim.x) {
        const double2 ls = B.Ls[i], lu = B.Lu[i], rs = B.Rs[i], ru = B.Ru[i], nm = B.nrm[i];
        const double n1 = nm.x, n2 = nm.y;
        const double unL = n1 * lu.x + n2 * lu.y, utL = -n2 * lu.x + n1 * lu.y;
        const double unR = n1 * ru.x + n2 * ru.y, utR = -n2 * ru.x + n1 * ru.y;
        double smax = 0.0, pq = 0.0;
        FluxE F;
        if constexpr (KIND == GD2_FACE_HLL)
            F = hll_euler<true>(ls.x, unL, utL, ls.y, rs.x, unR, utR, rs.y, B.gamma, B.inv_g1, smax);
        else if constexpr (KIND == GD2_FACE_HLLC_ALEKSASHOV)
            F = hllc_euler<true>(ls.x, unL, utL, ls.y, rs.x, unR, utR, rs.y, B.gamma, B.inv_g1, smax);
        else if constexpr (KIND == GD2_FACE_GODUNOV)
            F = exact_euler(ls.x, unL, utL, ls.y, rs.x, unR, utR, rs.y, B.gamma, smax, B.diag);
        else if constexpr (KIND == GD2_FACE_HLLC_ALEKSASHOV2)
            F = hllc_aleksashov2_euler(ls.x, unL, utL, ls.y, rs.x, unR, utR, rs.y, B.gamma, B.inv_g1, smax);
        else if constexpr (KIND == GD2_FACE_HLLC_KOROLKOV)
            F = hllc_korolkov_euler<true>(ls.x, unL, utL, ls.y, rs.x, unR, utR, rs.y, B.gamma, B.inv_g1, smax);
        else if constexpr (KIND == GD2_FACE_HLLC_KOROLKOV_2D)
            F = hllc_toro_euler<true>(ls.x, unL, utL, ls.y, rs.x, unR, utR, rs.y, B.gamma, B.inv_g1, smax);
        else if constexpr (KIND == GD2_FACE_HLLCQ_KOROLKOV_2D)
            F = hllc_toro_euler<true, true>(ls.x, unL, utL, ls.y, rs.x, unR, utR, rs.y, B.gamma, B.inv_g1, smax,
                                            B.QL[i], B.QR[i], &pq);
        else
            F = hllc_euler<true, true>(ls.x, unL, utL, ls.y, rs.x, unR, utR, rs.y, B.gamma, B.inv_g1, smax,
                                       B.QL[i], B.QR[i], &pq);
        B.Ps[i] = make_double2(F.fr, F.fe);
        B.Pu[i] = make_double2(n1 * F.fn - n2 * F.ft, n2 * F.fn + n1 * F.ft);
        B.dt[i] = B.krit * B.rad / smax;
        if constexpr (KIND == GD2_FACE_HLLCQ_KOROLKOV_2D || KIND == GD2_FACE_HLLCQ_ALEKSASHOV) B.PQ[i] = pq;
    }
}

struct MhdFaceBatch {
    const double *L, *R;  // n x 8: rho, p, vx, vy, vz, Bx, By, Bz (Gaussian B)
    const double* nrm;    // n x 3: face normal
    double* P;            // n x 8: F_rho, F_mx, F_my, F_mz, F_Bx, F_By, F_Bz, F_E
    double* dt;
    long long n;
    double gamma, inv_g1, krit, rad;
};

// HLLDQ_Korolkov (Gas_Dinamic.cu:964-1510) without its passive scalar: the
// face frame of its normal table (Gas_Dinamic.cu:978-1019), fields in
// Heaviside-Lorentz units inside, the flux back in the lab frame with the
// field fluxes in Gaussian units (Gas_Dinamic.cu:1488-1506).
template <int METOD>
__global__ __launch_bounds__(256) void mhd_face_kernel(const MhdFaceBatch B) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < B.n;
         i += (long long)gridDim.x * blockDim.x) {
        const double* l = B.L + 8 * i;
        const double* r = B.R + 8 * i;
        const double n1 = B.nrm[3 * i], n2 = B.nrm[3 * i + 1], n3 = B.nrm[3 * i + 2];
        double t1 = 0, t2 = 0, t3 = 0, m1 = 0, m2 = 0, m3 = 0;
        if (n1 > 0.1) {
            t2 = 1.0;
            m3 = 1.0;
        } else if (n2 > 0.1) {
            t3 = 1.0;
            m1 = 1.0;
        } else if (n3 > 0.1) {
            t1 = 1.0;
            m2 = 1.0;
        } else if (n1 < -0.1) {
            t3 = -1.0;
            m2 = -1.0;
        } else if (n2 < -0.1) {
            t1 = -1.0;
            m3 = -1.0;
        } else if (n3 < -0.1) {
            t1 = -1.0;
            m2 = -1.0;
        }
        auto dotv = [](const double* v, double a, double b, double c) { return v[0] * a + v[1] * b + v[2] * c; };
        double bl[3], br[3];
        for (int k = 0; k < 3; ++k) {
            bl[k] = l[5 + k] / kSpi4;
            br[k] = r[5 + k] / kSpi4;
        }
        double smax = 0.0;
        const FluxM PO = hlld_face<METOD>(
            l[0], l[1], dotv(l + 2, n1, n2, n3), dotv(l + 2, t1, t2, t3), dotv(l + 2, m1, m2, m3),
            dotv(bl, n1, n2, n3), dotv(bl, t1, t2, t3), dotv(bl, m1, m2, m3), r[0], r[1], dotv(r + 2, n1, n2, n3),
            dotv(r + 2, t1, t2, t3), dotv(r + 2, m1, m2, m3), dotv(br, n1, n2, n3), dotv(br, t1, t2, t3),
            dotv(br, m1, m2, m3), B.gamma, B.inv_g1, smax);
        double* P = B.P + 8 * i;
        P[0] = PO.f[0];
        P[1] = n1 * PO.f[1] + t1 * PO.f[2] + m1 * PO.f[3];
        P[2] = n2 * PO.f[1] + t2 * PO.f[2] + m2 * PO.f[3];
        P[3] = n3 * PO.f[1] + t3 * PO.f[2] + m3 * PO.f[3];
        P[4] = kSpi4 * (n1 * PO.f[5] + t1 * PO.f[6] + m1 * PO.f[7]);
        P[5] = kSpi4 * (n2 * PO.f[5] + t2 * PO.f[6] + m2 * PO.f[7]);
        P[6] = kSpi4 * (n3 * PO.f[5] + t3 * PO.f[6] + m3 * PO.f[7]);
        P[7] = PO.f[4];
        B.dt[i] = B.krit * B.rad / smax;
    }
}

static unsigned face_blocks(long long n) {
    long long b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b < 1 ? 1 : b);
}

DeviceDiag* device_diag();

}  // namespace gd2

using namespace gd2;

extern "C" {

int gd2_euler_face_fluxes(int kind, double gamma, double krit, double rad, long long n, const gd2_double2* Ls,
                          const gd2_double2* Lu, const gd2_double2* Rs, const gd2_double2* Ru, const double* QL,
                          const double* QR, const gd2_double2* normal, gd2_double2* Ps, gd2_double2* Pu, double* PQ,
                          double* dt, void* stream) {
    if (n < 0 || !(gamma > 1.0)) return GD2_ERR_ARG;
    if (kind < GD2_FACE_HLL || kind > GD2_FACE_HLLCQ_ALEKSASHOV) return GD2_ERR_UNSUPPORTED;
    if (n == 0) return GD2_OK;
    if (!Ls || !Lu || !Rs || !Ru || !normal || !Ps || !Pu || !dt) return GD2_ERR_ARG;
    const bool q = kind == GD2_FACE_HLLCQ_KOROLKOV_2D || kind == GD2_FACE_HLLCQ_ALEKSASHOV;
    if (q && (!QL || !QR || !PQ)) return GD2_ERR_ARG;
    FaceBatch B;
    B.Ls = reinterpret_cast<const double2*>(Ls);
    B.Lu = reinterpret_cast<const double2*>(Lu);
    B.Rs = reinterpret_cast<const double2*>(Rs);
    B.Ru = reinterpret_cast<const double2*>(Ru);
    B.QL = QL;
    B.QR = QR;
    B.nrm = reinterpret_cast<const double2*>(normal);
    B.Ps = reinterpret_cast<double2*>(Ps);
    B.Pu = reinterpret_cast<double2*>(Pu);
    B.PQ = PQ;
    B.dt = dt;
    B.n = n;
    B.gamma = gamma;
    B.inv_g1 = 1.0 / (gamma - 1.0);
    B.krit = krit;
    B.rad = rad;
    B.diag = device_diag();
    if (!B.diag) return GD2_ERR_HIP;
    const hipStream_t st = (hipStream_t)stream;
    const dim3 grid(face_blocks(n)), block(256);
    switch (kind) {
        case GD2_FACE_HLL: hipLaunchKernelGGL(euler_face_kernel<GD2_FACE_HLL>, grid, block, 0, st, B); break;
        case GD2_FACE_HLLC_ALEKSASHOV:
            hipLaunchKernelGGL(euler_face_kernel<GD2_FACE_HLLC_ALEKSASHOV>, grid, block, 0, st, B);
            break;
        case GD2_FACE_GODUNOV: hipLaunchKernelGGL(euler_face_kernel<GD2_FACE_GODUNOV>, grid, block, 0, st, B); break;
        case GD2_FACE_HLLC_ALEKSASHOV2:
            hipLaunchKernelGGL(euler_face_kernel<GD2_FACE_HLLC_ALEKSASHOV2>, grid, block, 0, st, B);
            break;
        case GD2_FACE_HLLC_KOROLKOV:
            hipLaunchKernelGGL(euler_face_kernel<GD2_FACE_HLLC_KOROLKOV>, grid, block, 0, st, B);
            break;
        case GD2_FACE_HLLC_KOROLKOV_2D:
            hipLaunchKernelGGL(euler_face_kernel<GD2_FACE_HLLC_KOROLKOV_2D>, grid, block, 0, st, B);
            break;
        case GD2_FACE_HLLCQ_KOROLKOV_2D:
            hipLaunchKernelGGL(euler_face_kernel<GD2_FACE_HLLCQ_KOROLKOV_2D>, grid, block, 0, st, B);
            break;
        default: hipLaunchKernelGGL(euler_face_kernel<GD2_FACE_HLLCQ_ALEKSASHOV>, grid, block, 0, st, B); break;
    }
    return hip_error_code(hipGetLastError());
}

int gd2_kernel_gkr(const gd2_grid* g, const gd2_double2* s, const gd2_double2* u, const gd2_double3* b,
                   gd2_double2* s2, gd2_double2* u2, gd2_double3* b2, double* T, const double* T_do, int i, int method,
                   void* stream) {
    (void)i;
    if (!g || g->nx < 1 || g->ny < 1 || !(g->dx > 0.0) || !(g->dy > 0.0) || !(g->gamma > 1.0)) return GD2_ERR_ARG;
    if ((long long)g->nx * g->ny > (1ll << 40)) return GD2_ERR_ARG;
    // Kernel_GKR's sides are its own (kernel.cu:2697-2770): symmetry west,
    // outflow east, axis south, inflow north; its sources are axisymmetric
    if (g->bc[0] != GD2_BC_SYMMETRY || g->bc[1] != GD2_BC_OUTFLOW || g->bc[2] != GD2_BC_AXIS ||
        g->bc[3] != GD2_BC_INFLOW)
        return GD2_ERR_UNSUPPORTED;
    if (!g->axisym || !(g->y0 > 0.0) || !(g->r_frozen >= 0.0)) return GD2_ERR_ARG;
    if (!s || !u || !b || !s2 || !u2 || !b2 || !T || !T_do) return GD2_ERR_ARG;
    if ((const void*)s == (const void*)s2 || (const void*)u == (const void*)u2 || (const void*)b == (const void*)b2)
        return GD2_ERR_ARG;
    if (method != 2 && method != 3) return GD2_ERR_UNSUPPORTED;
    GkrArgs A;
    A.s = reinterpret_cast<const double2*>(s);
    A.u = reinterpret_cast<const double2*>(u);
    A.b = reinterpret_cast<const double3*>(b);
    A.s2 = reinterpret_cast<double2*>(s2);
    A.u2 = reinterpret_cast<double2*>(u2);
    A.b2 = reinterpret_cast<double3*>(b2);
    A.tmin_bits = reinterpret_cast<unsigned long long*>(T);
    A.T_do = T_do;
    A.nx = g->nx;
    A.ny = g->ny;
    A.dx = g->dx;
    A.dy = g->dy;
    A.x0 = g->x0;
    A.y0 = g->y0;
    A.gamma = g->gamma;
    A.inv_g1 = 1.0 / (g->gamma - 1.0);
    A.idV = 1.0 / (g->dx * g->dy);
    A.krit = g->krit;
    A.r_frozen = g->r_frozen;
    A.diag = device_diag();
    if (!A.diag) return GD2_ERR_HIP;
    const hipStream_t st = (hipStream_t)stream;
    // grids with an interior (nx, ny >= 5): one launch, the interior march's
    // strips and then one-wave blocks over the two-cell boundary band;
    // smaller grids per cell
    if (g->nx < 5 || g->ny < 5) {
        const unsigned blocks = (unsigned)(((long long)g->nx * g->ny + 255) / 256);
        if (method == 3)
            hipLaunchKernelGGL(gkr_kernel<3>, dim3(blocks), dim3(256), 0, st, A);
        else
            hipLaunchKernelGGL(gkr_kernel<2>, dim3(blocks), dim3(256), 0, st, A);
        return hip_error_code(hipGetLastError());
    }
    A.tiles_x = (g->nx - 4 + kGkrW - 1) / kGkrW;
    const void* k = method == 3 ? reinterpret_cast<const void*>(&gkr_march_kernel<3>)
                                : reinterpret_cast<const void*>(&gkr_march_kernel<2>);
    A.rows = strip_rows(k, A.tiles_x, g->ny - 4, 2);
    A.march_blocks = A.tiles_x * ((g->ny - 4 + A.rows - 1) / A.rows);
    const long long band = 4LL * g->nx + 4LL * (g->ny - 4);
    const unsigned blocks = (unsigned)(A.march_blocks + (band + kWave - 1) / kWave);
    if (method == 3)
        hipLaunchKernelGGL(gkr_march_kernel<3>, dim3(blocks), dim3(kWave), 0, st, A);
    else
        hipLaunchKernelGGL(gkr_march_kernel<2>, dim3(blocks), dim3(kWave), 0, st, A);
    return hip_error_code(hipGetLastError());
}

int gd2_mhd_face_fluxes(int metod, double gamma, double krit, double rad, long long n, const double* L,
                        const double* R, const double* normal, double* P, double* dt, void* stream) {
    if (n < 0 || !(gamma > 1.0)) return GD2_ERR_ARG;
    if (metod != 2 && metod != 3) return GD2_ERR_UNSUPPORTED;
    if (n == 0) return GD2_OK;
    if (!L || !R || !normal || !P || !dt) return GD2_ERR_ARG;
    MhdFaceBatch B;
    B.L = L;
    B.R = R;
    B.nrm = normal;
    B.P = P;
    B.dt = dt;
    B.n = n;
    B.gamma = gamma;
    B.inv_g1 = 1.0 / (gamma - 1.0);
    B.krit = krit;
    B.rad = rad;
    const hipStream_t st = (hipStream_t)stream;
    if (metod == 2)
        hipLaunchKernelGGL(mhd_face_kernel<2>, dim3(face_blocks(n)), dim3(256), 0, st, B);
    else
        hipLaunchKernelGGL(mhd_face_kernel<3>, dim3(face_blocks(n)), dim3(256), 0, st, B);
    return hip_error_code(hipGetLastError());
}

}  // extern "C"
