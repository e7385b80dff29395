// Internal launchers shared between the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "dtv_geom.h"

namespace dtv {

int launch_fwd_brick(const float* vol, float* sino, int batch, const Geo& g, int bx, int by,
                     int bz, int block, hipStream_t st);
int launch_adj_brick(const float* sino, float* vol, int batch, const Geo& g, int bx, int by,
                     int bz, int block, hipStream_t st);

}  // namespace dtv
