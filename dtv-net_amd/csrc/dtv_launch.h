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
#ifndef DTV_FWD_STY_PAD
#define DTV_FWD_STY_PAD 0
#endif
#ifndef DTV_FWD_STZ_RES
#define DTV_FWD_STZ_RES 17
#endif
#ifndef DTV_FWD_STZ_MOD
#define DTV_FWD_STZ_MOD 32
#endif
#ifndef DTV_ADJ_STY_PAD
#define DTV_ADJ_STY_PAD 1
#endif
#ifndef DTV_ADJ_STZ_PAD
#define DTV_ADJ_STZ_PAD 1
#endif
__host__ __device__ constexpr int fwd_tile_sty(int bx) { return bx + 1 + DTV_FWD_STY_PAD; }
__host__ __device__ constexpr int fwd_tile_stz(int bx, int by) {
  return fwd_tile_sty(bx) * (by + 1) +
         ((DTV_FWD_STZ_RES - fwd_tile_sty(bx) * (by + 1)) & (DTV_FWD_STZ_MOD - 1));
}
// A^T rows hold the brick's bx voxels, the halo on both sides and padding.
__host__ __device__ constexpr int adj_tile_sty(int bx) { return bx + 2 + DTV_ADJ_STY_PAD; }
__host__ __device__ constexpr int adj_tile_stz(int bx, int by) {
  return adj_tile_sty(bx) * (by + 2) + DTV_ADJ_STZ_PAD;
}
// Partitioned A^T tiles (each sample handled by the brick holding its floor
// cell): x, y and z from the brick's first voxel to the upper halo, rows of
// bx + 2 words (the last a pad word, which also takes the taps at x = -1 of
// the next row's floor -1 samples; a leading pad word takes the first row's),
// z stride odd; then one FP32 accumulator per upper-halo voxel.
__host__ __device__ constexpr int adj_ptile_sty(int bx) { return bx + 2; }
__host__ __device__ constexpr int adj_ptile_stz(int bx, int by) {
  return (adj_ptile_sty(bx) * (by + 1)) | 1;
}
// Upper-halo voxels of a brick: the tile voxels with lx == bx, ly == by or
// lz == bz (local coordinates 0..b per axis).
__host__ __device__ constexpr int adj_halo_count(int bx, int by, int bz) {
  return (bx + 1) * (by + 1) * (bz + 1) - bx * by * bz;
}
inline size_t fwd_tile_bytes(int bx, int by, int bz) {
  return (size_t)fwd_tile_stz(bx, by) * (bz + 1) * 4;
}
inline size_t adj_tile_bytes(int bx, int by, int bz) {
  return (size_t)adj_tile_stz(bx, by) * (bz + 2) * 4;
}
inline size_t adj_ptile_bytes(int bx, int by, int bz) {
  return (1 + (size_t)adj_ptile_stz(bx, by) * (bz + 1) + adj_halo_count(bx, by, bz)) * 4;
}
constexpr size_t kMaxTileBytes = 56 * 1024;
// A^T brick kernels: owned voxels per lane (bx*by*bz <= kAdjPerThread * lanes)
constexpr int kAdjPerThread = 32;

// DTV_OK when p is coarse-grained device memory (hipMalloc), which the
// kernels' hardware FP32 global atomics need; DTV_ERR_INVALID_ARGUMENT for
// managed, host or fine-grained memory and pointers HIP does not know.
int check_atomic_target(const void* p);

// parts: workgroups per brick, each taking an equal share of the brick's views
// (grid z).  A's parts add into the sinogram like whole bricks.  A^T's parts
// either add into an output the caller has zero-filled (ws == nullptr; with
// two parts the sum does not depend on their order, since FP32 addition is
// commutative) or store their partial volumes into ws (parts x batch x voxels
// floats), which a second kernel adds up in part order into vol.
// fix_ws (optional, zero-filled by the caller, 256 + 8 x rays bytes): add the
// ray sums in int64 fixed point and convert at the end (bitwise reproducible).
int launch_fwd_brick(const float* vol, float* sino, int batch, const Geo& g, int bx, int by,
                     int bz, int block, int parts, hipStream_t st, void* fix_ws = nullptr);
int launch_adj_brick(const float* sino, float* vol, float* ws, int batch, const Geo& g, int bx,
                     int by, int bz, int block, int parts, hipStream_t st);
// Partitioned A^T: every sample is marched by the one brick holding its floor
// cell; a brick's taps on its upper halo are summed per voxel in FP32 and
// stored in ws, and a second kernel adds them into the neighbours' voxels in
// one fixed order (bitwise reproducible).  ws: adj_part_ws_bytes (every
// part's halo sums, then the partial volumes of parts 1..).  into (at most
// two parts): the parts' owned sums go straight into vol, which the caller
// zero-filled for two (added with FP32 atomics: order-independent for two
// terms); else part 0 stores into vol, the others into ws, and the second
// kernel adds them in part order.
size_t adj_part_ws_bytes(int batch, const Geo& g, int bx, int by, int bz, int parts);
int launch_adj_part(const float* sino, float* vol, float* ws, int batch, const Geo& g, int bx,
                    int by, int bz, int block, int parts, bool into, hipStream_t st);

}  // namespace dtv
