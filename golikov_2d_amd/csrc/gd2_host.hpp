// gd2_host.hpp -- host-side helpers shared by the translation units of
// libgd2.so (defined in gd2_api.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace gd2 {

// Rows per wave strip of the stage launches (GD2_TILE_ROWS overrides).
int tile_rows_for(int nx, int ny);
// Strip height for a launch of one-wave strips of `kernel`: the height that
// minimises rounds of resident waves x (rows + overhead).
int strip_rows(const void* kernel, int tiles_x, int ny, int overhead, int max_rows = 64, int min_rows = 4);
// GD2_OK, or GD2_ERR_HIP after recording the code for gd2_last_hip_error.
int hip_error_code(hipError_t e);

}  // namespace gd2
