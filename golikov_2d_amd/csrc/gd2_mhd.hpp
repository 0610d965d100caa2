// gd2_mhd.hpp -- ideal-MHD stage kernels (HLLC-MHD / HLLD + Powell source).
#pragma once

#include "gd2_device.hpp"

namespace gd2 {

inline int mhd_stage(const gd2_grid*, const gd2_scheme*, const double*, const double*, double*, double, double,
                     double, const double*, unsigned long long*, DeviceDiag*, hipStream_t) {
    return GD2_ERR_UNSUPPORTED;
}

inline int mhd_wavespeed(const gd2_grid*, const double*, unsigned long long*, hipStream_t) {
    return GD2_ERR_UNSUPPORTED;
}

inline int mhd_ref_step(const gd2_grid*, int, const gd2_double2*, const gd2_double2*, const gd2_double3*,
                        gd2_double2*, gd2_double2*, gd2_double3*, double*, const double*, int, DeviceDiag*,
                        hipStream_t) {
    return GD2_ERR_UNSUPPORTED;
}

}  // namespace gd2
