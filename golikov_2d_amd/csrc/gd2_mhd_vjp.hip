// gd2_mhd_vjp.hip -- the MHD stage VJP launches (face march, gather march,
// fold; the kernels are in gd2_mhd_adjoint.hpp). A translation unit of its
// own so that these kernels take their own code-generation options: the
// machine scheduler's ILP strategy (build.py), which cuts the HLLD face
// march's spill from 48 to 32 bytes and its stage VJP by about 4 %, where
// the same strategy slows the Euler kernels.
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/gd2.h"
#include "gd2_host.hpp"
#include "gd2_mhd_adjoint.hpp"

namespace gd2 {

template <int ORDER, int METOD>
static int launch_mhd_vjp(MhdAdjArgs J, hipStream_t st) {
    {
        MhdAdjArgs F = J;
        F.tiles_x = (F.nx + kFaceW - 1) / kFaceW;
        F.rows = strip_rows(reinterpret_cast<const void*>(&mhd_vjp_face_march_kernel<ORDER, METOD>), F.tiles_x, F.ny,
                            1, 64, 1);
        const unsigned blocks = (unsigned)(F.tiles_x * ((F.ny + F.rows - 1) / F.rows));
        hipLaunchKernelGGL((mhd_vjp_face_march_kernel<ORDER, METOD>), dim3(blocks), dim3(kWave), 0, st, F);
        const int e = hip_error_code(hipGetLastError());
        if (e != GD2_OK) return e;
    }
    J.tiles_x = (J.nx + kGatherW - 1) / kGatherW;
    J.rows = strip_rows(reinterpret_cast<const void*>(&mhd_vjp_gather_kernel<ORDER>), J.tiles_x, J.ny, 4);
    const unsigned blocks = (unsigned)(J.tiles_x * ((J.ny + J.rows - 1) / J.rows));
    hipLaunchKernelGGL((mhd_vjp_gather_kernel<ORDER>), dim3(blocks), dim3(kWave), 0, st, J);
    const int e = hip_error_code(hipGetLastError());
    if (e != GD2_OK) return e;
    // the boundary band: up to 4 columns x ny + 4 rows x nx cells
    hipLaunchKernelGGL(mhd_vjp_fold_kernel, dim3((unsigned)((4LL * (J.nx + J.ny) + kWave - 1) / kWave)), dim3(kWave), 0,
                       st, J);
    return hip_error_code(hipGetLastError());
}

int mhd_stage_vjp(const gd2_grid* g, const gd2_scheme* sc, const double* U, const double* V, const double* A,
                         double a, const double* B, double b, double c, double* W, void* workspace,
                         hipStream_t st) {
    MhdAdjArgs J;
    std::memset(&J, 0, sizeof J);
    J.U = U;
    J.V = V;
    J.A = A;
    J.B = B;
    J.W = W;
    J.a = a;
    J.b = b;
    J.c = c;
    J.nx = g->nx;
    J.ny = g->ny;
    J.plane = (long long)g->nx * g->ny;
    J.nfy = (long long)g->nx * (g->ny + 1);
    J.cx = static_cast<double*>(workspace);
    J.wy = J.cx + 7 * J.plane;
    J.gx = J.wy + 14 * J.nfy;
    J.gy = J.gx + 28LL * g->ny;
    J.dx = g->dx;
    J.dy = g->dy;
    J.dV = g->dx * g->dy;
    J.idV = 1.0 / J.dV;
    J.hfx = g->dy / J.dV;
    J.hfy = g->dx / J.dV;
    J.gamma = g->gamma;
    J.g1 = g->gamma - 1.0;
    J.inv_g1 = 1.0 / (g->gamma - 1.0);
    for (int s2 = 0; s2 < 4; ++s2) J.bc[s2] = g->bc[s2];
    const bool hlld = sc->flux == GD2_FLUX_HLLD;
    if (sc->order == 2) return hlld ? launch_mhd_vjp<2, 3>(J, st) : launch_mhd_vjp<2, 2>(J, st);
    return hlld ? launch_mhd_vjp<1, 3>(J, st) : launch_mhd_vjp<1, 2>(J, st);
}

}  // namespace gd2
