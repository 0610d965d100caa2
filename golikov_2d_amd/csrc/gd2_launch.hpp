// gd2_launch.hpp -- launchers of the Euler stage kernel and the one-pass
// SSP-RK2 step, shared by the translation units that instantiate them:
// gd2_api.hip (HLLC, exact, HLL) and gd2_variants.hip (the HLLC variants).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gd2.h"
#include "gd2_euler.hpp"
#include "gd2_host.hpp"
#include "gd2_rk2.hpp"

namespace gd2 {

inline unsigned euler_blocks(const EulerArgs& A) {
    return (unsigned)(A.tiles_x * ((A.ny + A.rows - 1) / A.rows));
}

template <int ORDER, int FLUX, int LAYOUT, bool W, bool BASE, int PROD = kRefPlain>
int launch_euler(const EulerArgs& A0, hipStream_t st) {
    EulerArgs A = A0;
    const void* k = reinterpret_cast<const void*>(&euler_tile_kernel<ORDER, FLUX, LAYOUT, W, BASE, PROD>);
    // strips down to one row: on grids too small to fill the chip each wave's
    // latency-bound march is the step's critical path
    A.rows = strip_rows(k, A.tiles_x, A.ny, 2, 64, 1);
    hipLaunchKernelGGL((euler_tile_kernel<ORDER, FLUX, LAYOUT, W, BASE, PROD>), dim3(euler_blocks(A)), dim3(kWave), 0,
                       st, A);
    return hip_error_code(hipGetLastError());
}

// One-pass SSP-RK2 step: strips of 59 columns; the strip height balances
// rounds of resident waves against the extra passes of each strip (four U1
// rows below and above it and the lead-in, about three rows of work), up to
// 256 rows so that a large grid fits in one round.
template <int ORDER, int FLUX, bool KEEP>
int launch_rk2(const EulerArgs& A0, hipStream_t st) {
    EulerArgs A = A0;
    A.tiles_x = (A.nx + kRk2W - 1) / kRk2W;
    const void* k = reinterpret_cast<const void*>(&euler_rk2_kernel<ORDER, FLUX, KEEP>);
    A.rows = max(2, strip_rows(k, A.tiles_x, A.ny, 3, 256));  // strips start at row 0 or row >= 2
    const unsigned blocks = (unsigned)(A.tiles_x * ((A.ny + A.rows - 1) / A.rows));
    hipLaunchKernelGGL((euler_rk2_kernel<ORDER, FLUX, KEEP>), dim3(blocks), dim3(kWave), 0, st, A);
    return hip_error_code(hipGetLastError());
}

// The stage and one-pass launches of the HLLC variants (GD2_FLUX_HLLC_KOROLKOV,
// GD2_FLUX_HLLC_2D), compiled in gd2_variants.hip; prod: kRefPlain /
// kRefOptions / kRefWestInflow. GD2_ERR_UNSUPPORTED for other fluxes or orders.
int launch_euler_variant(int order, int flux, int layout, bool w, bool base, int prod, const EulerArgs& A,
                         hipStream_t st);
int launch_rk2_variant(int order, int flux, bool keep, const EulerArgs& A, hipStream_t st);

inline bool is_hllc_variant(int flux) { return flux == GD2_FLUX_HLLC_KOROLKOV || flux == GD2_FLUX_HLLC_2D; }

}  // namespace gd2
