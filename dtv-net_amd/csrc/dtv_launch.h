// Internal launchers shared between the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "dtv_geom.h"

namespace dtv {

// LDS of the brick kernels' tiles: A holds the brick's voxels [o, o+b]
// (FP32, z stride padded), A^T its owned voxels plus a one-voxel halo as int32 with odd row
// strides.  Both stay below kMaxTileBytes so that, with the per-workgroup
// scheduling state (~5 KiB), a workgroup fits in 64 KiB of LDS.
// A's z stride is padded to 17 (mod 32) so that rays one detector row apart,
// which sit one z plane apart in the tile, fall on far-apart banks.
__host__ __device__ inline int fwd_tile_stz(int bx, int by) {
  const int p = (bx + 1) * (by + 1);
  return p + ((17 - p) & 31);
}
inline size_t fwd_tile_bytes(int bx, int by, int bz) {
  return (size_t)fwd_tile_stz(bx, by) * (bz + 1) * 4;
}
inline size_t adj_tile_bytes(int bx, int by, int bz) {
  return (size_t)((bx + 3) * (by + 2) + 1) * (bz + 2) * 4;
}
constexpr size_t kMaxTileBytes = 56 * 1024;

// DTV_OK when p is coarse-grained device memory (hipMalloc), which the
// kernels' hardware FP32 global atomics need; DTV_ERR_INVALID_ARGUMENT for
// managed, host or fine-grained memory and pointers HIP does not know.
int check_atomic_target(const void* p);

// parts: workgroups per brick, each taking an equal share of the brick's views
// (grid z).  A's parts add into the sinogram like whole bricks; A^T's parts add
// into an output the caller has zero-filled (with two parts the sum does not
// depend on their order, since FP32 addition is commutative).
int launch_fwd_brick(const float* vol, float* sino, int batch, const Geo& g, int bx, int by,
                     int bz, int block, int parts, hipStream_t st);
int launch_adj_brick(const float* sino, float* vol, int batch, const Geo& g, int bx, int by,
                     int bz, int block, int parts, hipStream_t st);

}  // namespace dtv
