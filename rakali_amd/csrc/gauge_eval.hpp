// gauge_eval.hpp -- bilinear interpolation of a station (gauges_write,
// rki_gauges.F90:516-550), shared by the gauge recorder and the adjoint's
// gauge misfit.  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

namespace rk {

// One field at the station: the four cells around it in the reference's
// operation order.  (c1, j1) is the upper neighbour of (c0, j0), or
// (c0, j0) itself at the east / north edge, where the reference reads a
// ghost cell at weight 0 (the product is a zero either way).
__device__ __forceinline__ double station_bilinear(const double* f, long long pitch, int c0, int c1, int j0, int j1,
                                                   double wx, double wy) {
    return (1.0 - wx) * (1.0 - wy) * f[(long long)j0 * pitch + c0] + wx * (1.0 - wy) * f[(long long)j0 * pitch + c1]
           + (1.0 - wx) * wy * f[(long long)j1 * pitch + c0] + wx * wy * f[(long long)j1 * pitch + c1];
}

// eta = h + b at a station, each interpolated (gauges_write :541).
__device__ __forceinline__ double station_eta(const double* h, const double* bed, long long pitch, int nx, int ny,
                                              int c0, int j0, double wx, double wy) {
    const int c1 = min(c0 + 1, nx - 1), j1 = min(j0 + 1, ny - 1);
    return station_bilinear(h, pitch, c0, c1, j0, j1, wx, wy) + station_bilinear(bed, pitch, c0, c1, j0, j1, wx, wy);
}

}  // namespace rk
