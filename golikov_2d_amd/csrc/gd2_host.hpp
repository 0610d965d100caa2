// gd2_host.hpp -- host-side helpers shared by the translation units of
// libgd2.so (defined in gd2_api.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gd2.h"

namespace gd2 {

// Rows per wave strip of the stage launches (GD2_TILE_ROWS overrides).
int tile_rows_for(int nx, int ny);
// Strip height for a launch of one-wave strips of `kernel`: the height that
// minimises rounds of resident waves x (rows + overhead).
int strip_rows(const void* kernel, int tiles_x, int ny, int overhead, int max_rows = 64, int min_rows = 4);
// GD2_OK, or GD2_ERR_HIP after recording the code for gd2_last_hip_error.
int hip_error_code(hipError_t e);
// The MHD stage VJP W = a A + b B + c J_L(U)^T V (HLLC-MHD or HLLD, order 1
// or 2) with its workspace (defined in gd2_mhd_vjp.hip); arguments checked
// by the caller.
int mhd_stage_vjp(const gd2_grid* g, const gd2_scheme* sc, const double* U, const double* V, const double* A,
                  double a, const double* B, double b, double c, double* W, void* workspace, hipStream_t st);

}  // namespace gd2
