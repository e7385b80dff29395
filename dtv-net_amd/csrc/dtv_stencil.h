// Forward-difference operators D_x, D_y, D_z on a volume [Z][Y][X] (x
// fastest) and the launch shape shared by the fused TV and shrinkage kernels.
//
// Two boundary modes:
//   neumann   (0): s = (1, X, X*Y); d_a(i) = f(i + s_a) - f(i), and 0 on the
//                  far face of axis a;
//   reference (1): the sparse matrices of model/PDHG/Utils/Gradients.py:8-47:
//                  s = (1, Y, X*Y) -- the reference's D_y strides by Y -- with
//                  D_x carrying no +1 where (i + 1) % X == 0 and D_y / D_z
//                  none in their last s_a rows; a row without its +1 gives
//                  d = -f(i).
// D_a^T p at i is p_a(i - s_a) where row i - s_a has its +1 at column i,
// minus p_a(i) where row i carries its -1 (every row in reference mode, the
// rows off the far face in neumann mode).
//
// Launch shape: one lane per voxel, workgroups of 64 x 4 lanes over an x-y
// tile, blockIdx.x = tile (x tiles fastest), blockIdx.y = z, blockIdx.z =
// batch item, so (x, y, z) come from the block indices with one uniform
// division and the stencil's boundary tests are plain comparisons.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dtv {

constexpr int kTileX = 64, kTileY = 4;

struct Vol3 {
  int nx, ny, nz;
  int mode;
  int tiles_x;       // ceil(nx / kTileX)
  int rx;            // s_1 % nx: x of i - s_1 is x - rx (mod nx) in reference mode
  int64_t n, s1, s2;  // voxels, stride of D_y, stride of D_z
};

__host__ __device__ inline Vol3 make_vol3(int nx, int ny, int nz, int mode) {
  Vol3 v;
  v.nx = nx;
  v.ny = ny;
  v.nz = nz;
  v.mode = mode;
  v.tiles_x = (nx + kTileX - 1) / kTileX;
  v.n = (int64_t)nx * ny * nz;
  const int s1 = mode == 0 ? nx : ny;
  v.s1 = s1;
  v.s2 = (int64_t)nx * ny;
  v.rx = s1 % nx;
  return v;
}

inline dim3 vol3_grid(const Vol3& v, int batch) {
  return dim3((unsigned)(v.tiles_x * ((v.ny + kTileY - 1) / kTileY)), (unsigned)v.nz,
              (unsigned)batch);
}

struct Pos {
  int x, y, z;
  int64_t i;
};

// This lane's voxel; false outside the volume.
__device__ __forceinline__ bool vol3_pos(const Vol3& v, Pos& p) {
  const int ty = (int)blockIdx.x / v.tiles_x;
  const int tx = (int)blockIdx.x - ty * v.tiles_x;
  p.x = tx * kTileX + (int)(threadIdx.x & (kTileX - 1));
  p.y = ty * kTileY + (int)(threadIdx.x / kTileX);
  p.z = (int)blockIdx.y;
  p.i = ((int64_t)p.z * v.ny + p.y) * v.nx + p.x;
  return p.x < v.nx && p.y < v.ny;
}

__device__ __forceinline__ int64_t stride(const Vol3& v, int a) {
  return a == 0 ? 1 : (a == 1 ? v.s1 : v.s2);
}

// does row i of D_a carry its +1 (at column i + s_a)?
__device__ __forceinline__ bool has_next(const Vol3& v, const Pos& p, int a) {
  if (a == 0) return p.x != v.nx - 1;
  if (v.mode == 0) return a == 1 ? p.y != v.ny - 1 : p.z != v.nz - 1;
  return p.i < v.n - stride(v, a);
}

// does row i - s_a of D_a carry a +1 at column i?
__device__ __forceinline__ bool has_prev(const Vol3& v, const Pos& p, int a) {
  if (a == 0) return p.x != 0;
  if (v.mode == 0) return a == 1 ? p.y != 0 : p.z != 0;
  return p.i >= stride(v, a);
}

// does row i of D_a carry its -1?
__device__ __forceinline__ bool has_self(const Vol3& v, const Pos& p, int a) {
  return v.mode != 0 || has_next(v, p, a);
}

// position of voxel i - s_a (only the fields the tests above read for it)
__device__ __forceinline__ Pos back(const Vol3& v, const Pos& p, int a) {
  Pos q = p;
  q.i = p.i - stride(v, a);
  if (a == 0) {
    q.x = p.x - 1;
  } else if (v.mode == 0) {
    if (a == 1) q.y = p.y - 1;
    else q.z = p.z - 1;
  } else if (a == 1) {
    q.x = p.x - v.rx;
    if (q.x < 0) q.x += v.nx;
  }
  return q;
}

// d_a at p of the volume f
__device__ __forceinline__ float diff(const Vol3& v, const float* __restrict__ f, const Pos& p,
                                      int a) {
  const float fi = f[p.i];
  if (has_next(v, p, a)) return f[p.i + stride(v, a)] - fi;
  return v.mode == 0 ? 0.f : -fi;
}

}  // namespace dtv
