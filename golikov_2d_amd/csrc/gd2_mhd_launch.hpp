// gd2_mhd_launch.hpp -- host launchers of the MHD stage kernel
// (gd2_mhd.hpp): the native stage, the wave-speed maximum and the
// reference-layout steps add2 / Kernel_TVD. Kept apart from the kernels so
// that a translation unit that needs only the MHD device functions (the
// stage VJP's) does not instantiate the stage kernels.
#pragma once

#include "gd2_mhd.hpp"

namespace gd2 {


inline void fill_mhd(MhdArgs& A, const gd2_grid* g, DeviceDiag* diag) {
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
    A.diag = diag;
    A.rows = tile_rows_for(g->nx, g->ny);
    A.tiles_x = (g->nx + kTileW - 1) / kTileW;
}

template <int ORDER, int METOD, int LAYOUT, bool W, bool PROD = false>
inline int launch_mhd(const MhdArgs& A0, hipStream_t st) {
    MhdArgs A = A0;
    A.rows = strip_rows(reinterpret_cast<const void*>(&mhd_tile_kernel<ORDER, METOD, LAYOUT, W, PROD>), A.tiles_x,
                        A.ny, 2, 64, 1);
    const unsigned blocks = (unsigned)(A.tiles_x * ((A.ny + A.rows - 1) / A.rows));
    hipLaunchKernelGGL((mhd_tile_kernel<ORDER, METOD, LAYOUT, W, PROD>), dim3(blocks), dim3(kWave), 0, st, A);
    return hip_error_code(hipGetLastError());
}

template <int LAYOUT, bool W, bool PROD = false>
inline int dispatch_mhd(int order, int metod, const MhdArgs& A, hipStream_t st) {
    if (order == 2) {
        if (metod == 3) return launch_mhd<2, 3, LAYOUT, W, PROD>(A, st);
        if (metod == 2) return launch_mhd<2, 2, LAYOUT, W, PROD>(A, st);
    } else if (order == 1) {
        if (metod == 3) return launch_mhd<1, 3, LAYOUT, W, PROD>(A, st);
        if (metod == 2) return launch_mhd<1, 2, LAYOUT, W, PROD>(A, st);
    }
    return GD2_ERR_UNSUPPORTED;
}

inline int mhd_stage(const gd2_grid* g, const gd2_scheme* sc, const double* Uin, const double* Ubase, double* Uout,
                     double a, double b, double dt, const double* dt_dev, unsigned long long* wmax_bits,
                     DeviceDiag* diag, hipStream_t st) {
    if (!diag) return GD2_ERR_HIP;
    MhdArgs A;
    fill_mhd(A, g, diag);
    A.Uin = Uin;
    A.Ubase = Ubase;
    A.Uout = Uout;
    A.a = a;
    A.b_coef = b;
    A.plain = (a == 0.0 && b == 1.0) ? 1 : 0;
    A.dt = dt;
    A.dt_dev = dt_dev;
    A.wmax_bits = wmax_bits;
    const int metod = sc->flux == GD2_FLUX_HLLD ? 3 : 2;
    if (wmax_bits) return dispatch_mhd<kSoA, true>(sc->order, metod, A, st);
    return dispatch_mhd<kSoA, false>(sc->order, metod, A, st);
}

inline unsigned mhd_ew_blocks(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b > 2048 ? 2048 : (b < 1 ? 1 : b));
}

inline int mhd_wavespeed(const gd2_grid* g, const double* U, unsigned long long* wmax_bits, hipStream_t st) {
    const long long n = (long long)g->nx * g->ny;
    hipLaunchKernelGGL(mhd_wavespeed_kernel, dim3(mhd_ew_blocks(n)), dim3(256), 0, st, U, n, g->gamma, g->gamma - 1.0,
                       1.0 / g->dx, 1.0 / g->dy, wmax_bits);
    return hip_error_code(hipGetLastError());
}

inline int mhd_ref_step(const gd2_grid* g, int order, const gd2_double2* s, const gd2_double2* u,
                        const gd2_double3* b, gd2_double2* s2, gd2_double2* u2, gd2_double3* b2, double* T,
                        const double* T_do, int method, DeviceDiag* diag, hipStream_t st) {
    if (!s || !u || !b || !s2 || !u2 || !b2 || !T || !T_do) return GD2_ERR_ARG;
    if ((const void*)s == (const void*)s2 || (const void*)u == (const void*)u2 || (const void*)b == (const void*)b2)
        return GD2_ERR_ARG;
    if (method != 2 && method != 3) return GD2_ERR_UNSUPPORTED;
    if (!diag) return GD2_ERR_HIP;
    MhdArgs A;
    fill_mhd(A, g, diag);
    A.s = reinterpret_cast<const double2*>(s);
    A.u = reinterpret_cast<const double2*>(u);
    A.b = reinterpret_cast<const double3*>(b);
    A.s2 = reinterpret_cast<double2*>(s2);
    A.u2 = reinterpret_cast<double2*>(u2);
    A.b2 = reinterpret_cast<double3*>(b2);
    A.dt_dev = T_do;
    A.tmin_bits = reinterpret_cast<unsigned long long*>(T);
    A.plain = 1;
    A.a = 0.0;
    A.b_coef = 1.0;
    A.x0 = g->x0;
    A.y0 = g->y0;
    A.r_frozen = g->r_frozen;
    A.inv_gamma = 1.0 / g->gamma;
    A.axisym = g->axisym ? 1 : 0;
    A.inflow_b_zero = order == 2 ? 1 : 0;
    bool prod = g->axisym || g->r_frozen > 0.0;
    for (int s = 0; s < 4; ++s) prod = prod || g->bc[s] >= GD2_BC_SYMMETRY;
    if (prod) return dispatch_mhd<kAoS, false, true>(order, method, A, st);
    return dispatch_mhd<kAoS, false>(order, method, A, st);
}

}  // namespace gd2
