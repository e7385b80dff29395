// gfx950 kernels of the cone-beam projector pair and their C ABI.
//
//   k_fwd_gather    A: one lane per detector ray, explicit 8-tap FP32
//                   trilinear gathers from HBM (no texture unit), 32.32
//                   fixed-point sample positions, clipped to the samples
//                   whose support touches the volume.
//   k_adj_scatter   matched A^T: the same enumeration, each sample's 8
//                   weights scattered with hardware FP32 atomics.
//   k_bp_standard   reference backprojector of Standard/plug/kernel/
//                   backwardKernel.cu:22-53 (voxel-driven, bilinear, FDK
//                   weight), all views accumulated in registers.
//   k_bp_beijing    reference backprojector of Beijing/plug/kernel/
//                   backwardKernel.cu:18-45.
#include <hip/hip_runtime.h>

#include "dtv_geom.h"
#include "dtv_launch.h"
#include "dtvproj.h"

using namespace dtv;

namespace {

constexpr int kRayTileU = 64;   // rays per wave along u
constexpr int kRayTileV = 4;    // waves per block along v

// Trilinear sample at the fixed-point position (px, py, pz); taps outside
// the volume read as zero (border addressing of the reference texture).
__device__ __forceinline__ float tri_checked(const float* __restrict__ vol, int nx, int ny,
                                             int nz, int64_t px, int64_t py, int64_t pz) {
  const int ix = (int)(px >> 32), iy = (int)(py >> 32), iz = (int)(pz >> 32);
  const float fx = (float)(uint32_t)px * kInvFix;
  const float fy = (float)(uint32_t)py * kInvFix;
  const float fz = (float)(uint32_t)pz * kInvFix;
  const bool x0 = (unsigned)ix < (unsigned)nx, x1 = (unsigned)(ix + 1) < (unsigned)nx;
  const bool y0 = (unsigned)iy < (unsigned)ny, y1 = (unsigned)(iy + 1) < (unsigned)ny;
  const bool z0 = (unsigned)iz < (unsigned)nz, z1 = (unsigned)(iz + 1) < (unsigned)nz;
  const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
  const size_t base = (size_t)iz * sz + (size_t)iy * sy + (size_t)ix;
  float t[8];
  t[0] = (z0 && y0 && x0) ? vol[base] : 0.f;
  t[1] = (z0 && y0 && x1) ? vol[base + 1] : 0.f;
  t[2] = (z0 && y1 && x0) ? vol[base + sy] : 0.f;
  t[3] = (z0 && y1 && x1) ? vol[base + sy + 1] : 0.f;
  t[4] = (z1 && y0 && x0) ? vol[base + sz] : 0.f;
  t[5] = (z1 && y0 && x1) ? vol[base + sz + 1] : 0.f;
  t[6] = (z1 && y1 && x0) ? vol[base + sz + sy] : 0.f;
  t[7] = (z1 && y1 && x1) ? vol[base + sz + sy + 1] : 0.f;
  const float a0 = t[0] + fx * (t[1] - t[0]);
  const float a1 = t[2] + fx * (t[3] - t[2]);
  const float a2 = t[4] + fx * (t[5] - t[4]);
  const float a3 = t[6] + fx * (t[7] - t[6]);
  const float b0 = a0 + fy * (a1 - a0);
  const float b1 = a2 + fy * (a3 - a2);
  return b0 + fz * (b1 - b0);
}

__global__ __launch_bounds__(kRayTileU * kRayTileV) void k_fwd_gather(
    const float* __restrict__ vol, float* __restrict__ sino, Geo g, int batch) {
  const int iu = blockIdx.x * kRayTileU + threadIdx.x;
  const int iv = blockIdx.y * kRayTileV + threadIdx.y;
  const int a = blockIdx.z;
  if (iu >= g.nu || iv >= g.nv) return;
  View v;
  load_view(g, a, v);
  Ray r;
  ray_setup(g, v, iu, iv, r);
  const size_t vstride = (size_t)g.nx * g.ny * g.nz;
  const float scale = (float)g.scale;
  for (int b = 0; b < batch; ++b) {
    const float* vb = vol + (size_t)b * vstride;
    float acc = 0.f;
    int64_t px = r.P[0] + (int64_t)r.k0 * r.D[0];
    int64_t py = r.P[1] + (int64_t)r.k0 * r.D[1];
    int64_t pz = r.P[2] + (int64_t)r.k0 * r.D[2];
    for (int k = r.k0; k < r.k1; ++k) {
      acc += tri_checked(vb, g.nx, g.ny, g.nz, px, py, pz);
      px += r.D[0];
      py += r.D[1];
      pz += r.D[2];
    }
    sino[(((size_t)b * g.n_views + a) * g.nv + iv) * g.nu + iu] = acc * scale;
  }
}

__device__ __forceinline__ void scatter_tri(float* __restrict__ vol, int nx, int ny, int nz,
                                            int64_t px, int64_t py, int64_t pz, float val) {
  const int ix = (int)(px >> 32), iy = (int)(py >> 32), iz = (int)(pz >> 32);
  const float fx = (float)(uint32_t)px * kInvFix;
  const float fy = (float)(uint32_t)py * kInvFix;
  const float fz = (float)(uint32_t)pz * kInvFix;
  const bool x0 = (unsigned)ix < (unsigned)nx, x1 = (unsigned)(ix + 1) < (unsigned)nx;
  const bool y0 = (unsigned)iy < (unsigned)ny, y1 = (unsigned)(iy + 1) < (unsigned)ny;
  const bool z0 = (unsigned)iz < (unsigned)nz, z1 = (unsigned)(iz + 1) < (unsigned)nz;
  const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
  const size_t base = (size_t)iz * sz + (size_t)iy * sy + (size_t)ix;
  const float wz0 = (1.f - fz) * val, wz1 = fz * val;
  const float w00 = (1.f - fy) * wz0, w01 = fy * wz0, w10 = (1.f - fy) * wz1, w11 = fy * wz1;
  const float gx = 1.f - fx;
  if (z0 && y0 && x0) unsafeAtomicAdd(vol + base, gx * w00);
  if (z0 && y0 && x1) unsafeAtomicAdd(vol + base + 1, fx * w00);
  if (z0 && y1 && x0) unsafeAtomicAdd(vol + base + sy, gx * w01);
  if (z0 && y1 && x1) unsafeAtomicAdd(vol + base + sy + 1, fx * w01);
  if (z1 && y0 && x0) unsafeAtomicAdd(vol + base + sz, gx * w10);
  if (z1 && y0 && x1) unsafeAtomicAdd(vol + base + sz + 1, fx * w10);
  if (z1 && y1 && x0) unsafeAtomicAdd(vol + base + sz + sy, gx * w11);
  if (z1 && y1 && x1) unsafeAtomicAdd(vol + base + sz + sy + 1, fx * w11);
}

__global__ __launch_bounds__(kRayTileU * kRayTileV) void k_adj_scatter(
    const float* __restrict__ sino, float* __restrict__ vol, Geo g, int batch) {
  const int iu = blockIdx.x * kRayTileU + threadIdx.x;
  const int iv = blockIdx.y * kRayTileV + threadIdx.y;
  const int a = blockIdx.z;
  if (iu >= g.nu || iv >= g.nv) return;
  View v;
  load_view(g, a, v);
  Ray r;
  ray_setup(g, v, iu, iv, r);
  const size_t vstride = (size_t)g.nx * g.ny * g.nz;
  const float scale = (float)g.scale;
  for (int b = 0; b < batch; ++b) {
    const float val = sino[(((size_t)b * g.n_views + a) * g.nv + iv) * g.nu + iu] * scale;
    if (val == 0.f) continue;
    float* vb = vol + (size_t)b * vstride;
    int64_t px = r.P[0] + (int64_t)r.k0 * r.D[0];
    int64_t py = r.P[1] + (int64_t)r.k0 * r.D[1];
    int64_t pz = r.P[2] + (int64_t)r.k0 * r.D[2];
    for (int k = r.k0; k < r.k1; ++k) {
      scatter_tri(vb, g.nx, g.ny, g.nz, px, py, pz, val);
      px += r.D[0];
      py += r.D[1];
      pz += r.D[2];
    }
  }
}

__global__ void k_fill_zero(float* __restrict__ p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.f;
}

// Bilinear sample of one detector plane at a texel-centred coordinate given
// as integer part + fraction; zero outside (border addressing).
__device__ __forceinline__ float bil_checked(const float* __restrict__ img, int nu, int nv,
                                             int ix, float fx, int iy, float fy) {
  const bool x0 = (unsigned)ix < (unsigned)nu, x1 = (unsigned)(ix + 1) < (unsigned)nu;
  const bool y0 = (unsigned)iy < (unsigned)nv, y1 = (unsigned)(iy + 1) < (unsigned)nv;
  const size_t base = (size_t)iy * nu + ix;
  const float t00 = (y0 && x0) ? img[base] : 0.f;
  const float t01 = (y0 && x1) ? img[base + 1] : 0.f;
  const float t10 = (y1 && x0) ? img[base + nu] : 0.f;
  const float t11 = (y1 && x1) ? img[base + nu + 1] : 0.f;
  const float a0 = t00 + fx * (t01 - t00);
  const float a1 = t10 + fx * (t11 - t10);
  return a0 + fy * (a1 - a0);
}

constexpr int kBpTileX = 64;
constexpr int kBpZ = 16;          // voxels along z per lane
constexpr int kBpViewChunk = 64;  // views staged in LDS at a time

// Standard backwardKernel.cu:22-53.  For voxel centre c = ((ix-X/2)s, (iy-Y/2)s)
// and view a, the source-through-voxel line meets the detector line at
// u = SDD<c,ey>/L + U/2 with L = SID + <c,ex> (the closed form of
// helper_geometry.h:26-42's intersection), v_k = (k - Z/2) SDD s/L + V/2,
// weight (SDD/L)^2 pi/A.
__global__ __launch_bounds__(kBpTileX) void k_bp_standard(
    const float* __restrict__ sino, float* __restrict__ vol, Geo g, int b) {
  __shared__ double cs_[kBpViewChunk][2];
  const int ix = blockIdx.x * kBpTileX + threadIdx.x;
  const int iy = blockIdx.y;
  const int z0 = blockIdx.z * kBpZ;
  const double s = g.sid / g.sdd;
  const double cx = (ix - g.nx / 2.0) * s;
  const double cy = (iy - g.ny / 2.0) * s;
  const float wpi = (float)(kPiRef / g.n_views);
  // a view's byte offsets fit 32 bits and its rows 24-bit multiplies
  const bool offsets32 = (int64_t)g.nu * g.nv < (int64_t)1 << 29 && g.nu < (1 << 21);
  float acc[kBpZ];
#pragma unroll
  for (int k = 0; k < kBpZ; ++k) acc[k] = 0.f;
  const float* sb = sino + (size_t)b * g.n_views * g.nu * g.nv;
  for (int a0 = 0; a0 < g.n_views; a0 += kBpViewChunk) {
    // wave-uniform (a scalar register), so the view loop below is uniform
    const int na = __builtin_amdgcn_readfirstlane(min(kBpViewChunk, g.n_views - a0));
    __syncthreads();
    if (threadIdx.x < na) {
      const double th = g.angles[a0 + threadIdx.x] / 180.0 * kPiRef - kPiRef / 2.0;
      cs_[threadIdx.x][0] = cos(th);
      cs_[threadIdx.x][1] = sin(th);
    }
    __syncthreads();
    // lanes past the volume edge run too (their taps are bounds-checked and
    // their sums never written), so the view loop stays wave-uniform
    {
      for (int j = 0; j < na; ++j) {
        const double c = cs_[j][0], sn = cs_[j][1];
        const double L = g.sid + (cx * c + cy * sn);
        const double m = g.sdd / L;
        const double u = m * (cx * -sn + cy * c) + g.nu / 2.0;
        const double dz = m * s;
        const double fu = floor(u);
        const int iu = (int)fu;
        const float fx = (float)(u - fu);
        const float w = (float)(m * m) * wpi;
        const float* img = sb + (size_t)(a0 + j) * g.nu * g.nv;
        double vv = (z0 - g.nz / 2.0) * dz + g.nv / 2.0;
        // Fast path: the lane's kBpZ detector rows in 32.32 fixed point (row k
        // at vv + k dz, formed only where that cannot overflow; dz > 0, so
        // the rows increase with k and the first and last bound them all) and
        // every tap on the detector, read through 32-bit byte offsets from the
        // view's base.  Otherwise the bounds-checked gather of every tap.
        bool fast = false;
        int64_t V = 0, DV = 0;
        if (offsets32 && (unsigned)iu < (unsigned)(g.nu - 1) && dz > 0.0 && dz < 1048576.0 &&
            fabs(vv) < 536870912.0) {
          V = to_fix(vv);
          DV = to_fix(dz);
          fast = (V >> 32) >= 0 && ((V + (kBpZ - 1) * DV) >> 32) < g.nv - 1;
        }
        if (fast) {
          const char* base = reinterpret_cast<const char*>(img);
          const unsigned row = 4u * (unsigned)g.nu;
#pragma unroll
          for (int k = 0; k < kBpZ; ++k) {
            const unsigned o = __umul24((unsigned)(V >> 32), row) + 4u * (unsigned)iu;
            const float fy = (float)(uint32_t)V * kInvFix;
            const float t00 = *reinterpret_cast<const float*>(base + o);
            const float t01 = *reinterpret_cast<const float*>(base + o + 4u);
            const float t10 = *reinterpret_cast<const float*>(base + o + row);
            const float t11 = *reinterpret_cast<const float*>(base + o + row + 4u);
            const float a0 = t00 + fx * (t01 - t00);
            const float a1 = t10 + fx * (t11 - t10);
            acc[k] += w * (a0 + fy * (a1 - a0));
            V += DV;
          }
        } else {
#pragma unroll
          for (int k = 0; k < kBpZ; ++k) {
            const double fv = floor(vv);
            acc[k] += w * bil_checked(img, g.nu, g.nv, iu, fx, (int)fv, (float)(vv - fv));
            vv += dz;
          }
        }
      }
    }
  }
  if (ix >= g.nx) return;
  float* vb = vol + (size_t)b * g.nx * g.ny * g.nz;
#pragma unroll
  for (int k = 0; k < kBpZ; ++k) {
    const int iz = z0 + k;
    if (iz < g.nz) vb[((size_t)iz * g.ny + iy) * g.nx + ix] = acc[k];
  }
}

// Beijing backwardKernel.cu:18-45 with the triple products expanded per view:
// den = (S-c).(u x v), sx = ((c-S).(v x (S-D)))/den + U/2,
// sy = ((c-S).((S-D) x u))/den + V/2, fr = ((S-D).(u x v))/den; views whose
// (sx, sy) fall outside [-1, X+1] x [-1, Y+1] are skipped (the reference
// tests against the volume size); bilinear at (sx - 0.5, sy - 0.5).
struct BjView {
  double S[3], n[3], m1[3], m2[3], fr;
};

__global__ __launch_bounds__(kBpTileX) void k_bp_beijing(
    const float* __restrict__ sino, float* __restrict__ vol, Geo g, int b) {
  __shared__ BjView vw[kBpViewChunk];
  __shared__ int live[kBpViewChunk];
  const int ix = blockIdx.x * kBpTileX + threadIdx.x;
  const int iy = blockIdx.y;
  const int z0 = blockIdx.z * kBpZ;
  // this workgroup's voxels: x in [xa, xb], row iy, z in [z0, zb] (centres)
  const double xa = blockIdx.x * kBpTileX - g.nx / 2.0;
  const double xb = min((int)(blockIdx.x + 1) * kBpTileX, g.nx) - 1 - g.nx / 2.0;
  const double za = z0 - g.nz / 2.0, zb = min(z0 + kBpZ, g.nz) - 1 - g.nz / 2.0;
  float acc[kBpZ];
#pragma unroll
  for (int k = 0; k < kBpZ; ++k) acc[k] = 0.f;
  const float* sb = sino + (size_t)b * g.n_views * g.nu * g.nv;
  const double cx = ix - g.nx / 2.0, cy = iy - g.ny / 2.0;
  const double xlim = g.nx + 1.0, ylim = g.ny + 1.0;
  for (int a0 = 0; a0 < g.n_views; a0 += kBpViewChunk) {
    const int na = min(kBpViewChunk, g.n_views - a0);
    __syncthreads();
    if (threadIdx.x < na) {
      const float* p = g.pvec + 12 * (size_t)(a0 + threadIdx.x);
      const double S[3] = {(double)p[0], (double)p[1], -(double)p[2]};
      const double D[3] = {(double)p[3], (double)p[4], -(double)p[5]};
      const double vv[3] = {(double)p[6], (double)p[7], (double)p[8]};
      const double uu[3] = {(double)p[9], (double)p[10], (double)p[11]};
      const double sd[3] = {S[0] - D[0], S[1] - D[1], S[2] - D[2]};
      BjView& w = vw[threadIdx.x];
      w.n[0] = uu[1] * vv[2] - uu[2] * vv[1];
      w.n[1] = uu[2] * vv[0] - uu[0] * vv[2];
      w.n[2] = uu[0] * vv[1] - uu[1] * vv[0];
      w.m1[0] = vv[1] * sd[2] - vv[2] * sd[1];
      w.m1[1] = vv[2] * sd[0] - vv[0] * sd[2];
      w.m1[2] = vv[0] * sd[1] - vv[1] * sd[0];
      w.m2[0] = sd[1] * uu[2] - sd[2] * uu[1];
      w.m2[1] = sd[2] * uu[0] - sd[0] * uu[2];
      w.m2[2] = sd[0] * uu[1] - sd[1] * uu[0];
      w.fr = sd[0] * w.n[0] + sd[1] * w.n[1] + sd[2] * w.n[2];
      for (int c = 0; c < 3; ++c) w.S[c] = S[c];
      // Views whose detector this workgroup's voxels cannot reach add exact
      // zeros: when den has one strict sign at every corner of the voxel
      // rectangle (so on all of it: den is affine) the projective map sends
      // the rectangle inside the hull of the corner images, and a voxel adds
      // nothing once its taps (floor(s - 0.5), +1) all leave the detector,
      // i.e. sx < -0.5 or sx >= U + 0.5 (likewise sy); one more pixel of
      // slack covers rounding.
      int pos = 0, neg = 0;
      double sxl = 1e300, sxh = -1e300, syl = 1e300, syh = -1e300;
      for (int q = 0; q < 4; ++q) {
        const double c0 = ((q & 1) ? xb : xa) - S[0], c1 = (iy - g.ny / 2.0) - S[1];
        const double c2 = ((q & 2) ? zb : za) - S[2];
        const double den = -(c0 * w.n[0] + c1 * w.n[1] + c2 * w.n[2]);
        pos += den > 0.0;
        neg += den < 0.0;
        const double sx = (c0 * w.m1[0] + c1 * w.m1[1] + c2 * w.m1[2]) / den + g.nu / 2.0;
        const double sy = (c0 * w.m2[0] + c1 * w.m2[1] + c2 * w.m2[2]) / den + g.nv / 2.0;
        sxl = fmin(sxl, sx);
        sxh = fmax(sxh, sx);
        syl = fmin(syl, sy);
        syh = fmax(syh, sy);
      }
      const bool keep = pos != 4 && neg != 4;
      const bool out = sxh < -1.5 || sxl > g.nu + 1.5 || syh < -1.5 || syl > g.nv + 1.5;
      live[threadIdx.x] = keep || !out || !(sxl <= sxh && syl <= syh);
    }
    __syncthreads();
    if (ix < g.nx) {
      for (int j = 0; j < na; ++j) {
        if (!live[j]) continue;
        const BjView& w = vw[j];
        const double csx = cx - w.S[0], csy = cy - w.S[1];
        const double csz = (z0 - g.nz / 2.0) - w.S[2];
        const float* img = sb + (size_t)(a0 + j) * g.nu * g.nv;
        // the three triple products are affine in c: step them along z
        // (one add each per voxel; 16 steps add ~16 ulp to exact FP64 values)
        double den = -(csx * w.n[0] + csy * w.n[1] + csz * w.n[2]);
        double nu1 = csx * w.m1[0] + csy * w.m1[1] + csz * w.m1[2];
        double nu2 = csx * w.m2[0] + csy * w.m2[1] + csz * w.m2[2];
#pragma unroll
        for (int k = 0; k < kBpZ; ++k) {
          // reciprocal from the hardware estimate and two Newton steps
          // (about one ulp); a zero denominator keeps IEEE 1/0
          const double f = den == 0.0 ? copysign(__builtin_inf(), den) : recip(den);
          const double sx = f * nu1 + g.nu / 2.0;
          const double sy = f * nu2 + g.nv / 2.0;
          if (!(sx < -1.0 || sx > xlim || sy < -1.0 || sy > ylim)) {
            const double x = sx - 0.5, y = sy - 0.5;
            const double fxd = floor(x), fyd = floor(y);
            acc[k] += (float)(f * w.fr) *
                      bil_checked(img, g.nu, g.nv, (int)fxd, (float)(x - fxd), (int)fyd,
                                  (float)(y - fyd));
          }
          den -= w.n[2];
          nu1 += w.m1[2];
          nu2 += w.m2[2];
        }
      }
    }
  }
  if (ix >= g.nx) return;
  const float wgt = (float)(2.0 * kPiRef / (double)(g.n_views / 21));
  float* vb = vol + (size_t)b * g.nx * g.ny * g.nz;
#pragma unroll
  for (int k = 0; k < kBpZ; ++k) {
    const int iz = z0 + k;
    if (iz < g.nz) vb[((size_t)iz * g.ny + iy) * g.nx + ix] = acc[k] * wgt;
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

int check_sizes(int32_t batch, const int32_t* vs, const int32_t* ds, int32_t n_views) {
  if (!vs || !ds) return DTV_ERR_NULL_POINTER;
  if (batch < 1 || n_views < 1) return DTV_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 3; ++i)
    if (vs[i] < 1) return DTV_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 2; ++i)
    if (ds[i] < 1) return DTV_ERR_INVALID_ARGUMENT;
  const int64_t nvox = (int64_t)vs[0] * vs[1] * vs[2];
  const int64_t nray = (int64_t)ds[0] * ds[1];
  if (nvox >= (1ll << 31) || nray >= (1ll << 31)) return DTV_ERR_TOO_LARGE;
  if (n_views > 65535) return DTV_ERR_TOO_LARGE;
  if (vs[0] > (1 << 20) || vs[1] > (1 << 20) || vs[2] > (1 << 20)) return DTV_ERR_TOO_LARGE;
  return DTV_OK;
}

struct DeviceGuard {
  int prev = -1;
  bool changed = false;
  int status = DTV_OK;
  explicit DeviceGuard(int32_t device) {
    if (device < 0) return;
    if (hipGetDevice(&prev) != hipSuccess) {
      status = DTV_ERR_HIP;
      return;
    }
    if (prev != device) {
      if (hipSetDevice(device) != hipSuccess) {
        status = DTV_ERR_HIP;
        return;
      }
      changed = true;
    }
  }
  ~DeviceGuard() {
    if (changed) (void)hipSetDevice(prev);
  }
};

Geo make_geo(const int32_t* vs, const int32_t* ds, int32_t n_views) {
  Geo g{};
  g.nx = vs[0];
  g.ny = vs[1];
  g.nz = vs[2];
  g.nu = ds[0];
  g.nv = ds[1];
  g.n_views = n_views;
  return g;
}

void scheme_standard(Geo& g, double sid, double sdd) {
  const double s = sid / sdd;
  g.axis_mode = kAxisXY;
  const double n3[3] = {(double)g.nx, (double)g.ny, (double)g.nz};
  for (int c = 0; c < 3; ++c) {
    const double half = n3[c] / 2.0 / s;
    g.lo[c] = n3[c] / 2.0 - half;
    g.hi[c] = n3[c] / 2.0 + half;
  }
  g.margin = 3.0 / s;
  g.scale = s;
  g.sid = sid;
  g.sdd = sdd;
}

void scheme_beijing(Geo& g) {
  g.axis_mode = kAxisXYZ;
  g.lo[0] = g.lo[1] = g.lo[2] = 0.0;
  g.hi[0] = g.nx;
  g.hi[1] = g.ny;
  g.hi[2] = g.nz;
  g.margin = 3.0;
  g.scale = 1.0;
}

int launch_status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

int run_zero(float* p, size_t n, hipStream_t st) {
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, st, p, n);
  return launch_status();
}

// Resolved kernel choice.
struct Launch {
  int algo = DTV_ALGO_BRICK;
  int bx = 26, by = 26, bz = 14;
  int block = 256;
  int parts = 1;  // workgroups sharing one brick's views
  bool parts_requested = false;  // set by the caller (else a default)
};

// Bricks of shape (bx, by, bz) covering a volume: A's bricks tile floor(p)
// over [-1, N-1] (one more voxel per axis), A^T's the voxels.
long long brick_count(const Geo& g, int bx, int by, int bz, bool adjoint) {
  const int e = adjoint ? 0 : 1;
  return (long long)((g.nx + e + bx - 1) / bx) * ((g.ny + e + by - 1) / by) *
         ((g.nz + e + bz - 1) / bz);
}

// Default kernels, by problem size (measured on MI355X, 256 CUs, dense inputs):
//  * A: 26x26x14 bricks (45 KB of tile: three workgroups per CU, the most a
//    forward tile can hold at that occupancy; with the views split it beat
//    24x24x14 at 256^3 / 60 views 2.47 vs 2.60 ms, 128^3 / 90 views 0.92 vs
//    1.01 ms, 512x512x384 / 120 views 27.1 vs 28.8 ms and the 3024-view
//    Beijing scan 6.40 vs 6.64 ms; odd row strides only: 25x25x14 lost to
//    both) with their views split over parts = ceil(4608 / (bricks x batch))
//    workgroups per brick, at most 16 (about six rounds of the 768 workgroups
//    the chip holds at once): 256^3 / 60 views 2.62 -> 2.56 ms (3 parts, on
//    24x24x14), 128^3 / 90 views 1.79 ms (16x16x16, 1 part) -> 0.96 ms (13),
//    64^3 / 72 views 0.61 ms (ray kernel) -> 0.21 ms (16), the Beijing scan
//    7.75 -> 6.59 ms (8).  Only when even 16 parts would give under 256
//    workgroups, the one-lane-per-ray kernel (32^3 / 24 views: 0.042 ms
//    against 0.050 on bricks).
//  * A^T, through the workspace entry point (each part stores its partial
//    volumes, a second kernel adds them in part order: bitwise reproducible
//    for any split): 32x16x16 bricks (16x16x16 when fewer than 128 of those,
//    16x16x8 below 32 of 16^3), parts = ceil(4096 / (bricks x batch)), at most
//    16 (N(0,1) sinograms: 256^3 / 60 views 3.92 ms with 2 parts against 4.01
//    unsplit, 160^3 / 60 views 2.20 ms (9 parts) against 3.51, 128^3 / 90
//    views 1.87 ms (16) against 3.58, 64^3 / 72 views 0.37 ms (16, 16^3)
//    against 1.12 on 16x16x8, the Beijing scan 15.6 ms (8) against 17.4; at
//    512x512x384 / 120 views any split was slower, 46.0 unsplit against 50.2
//    ms).  Without a workspace (the reference-shaped entry points) at most two
//    parts can add into a zero-filled output (FP32 addition is commutative,
//    so the sum stays bitwise reproducible): two below 4096 bricks, on
//    32x16x16 bricks when they number at least 512 with the batch, else
//    16x16x16 when those number at least 200, else 16x16x8 (128^3: 2.33 vs
//    2.61 ms, 64^3: 0.59 vs 0.91 ms).  Outputs the hardware atomics cannot
//    address (managed or host memory) keep one part.
#ifndef DTV_ADJ_TARGET_GROUPS
#define DTV_ADJ_TARGET_GROUPS 4096
#endif
constexpr long long kFwdTargetGroups = 4608;
constexpr int kFwdMaxParts = 16;

constexpr long long kAdjTargetGroups = DTV_ADJ_TARGET_GROUPS;
constexpr int kAdjMaxParts = 16;      // with a workspace
constexpr int kAdjMaxAtomicParts = 2;  // adding into the output

// Parts per brick for a brick shape when the caller does not fix them.  A^T
// without a workspace takes two (bitwise reproducible atomic adds); with one,
// enough parts for about kAdjTargetGroups workgroups (at least two).
int default_parts(bool adjoint, const Geo& g, int batch, int bx, int by, int bz,
                  bool ws = false) {
  const long long nb = brick_count(g, bx, by, bz, adjoint) * batch;
  if (adjoint && !ws) return nb < kAdjTargetGroups ? kAdjMaxAtomicParts : 1;
  if (adjoint)
    return (int)std::max<long long>(
        1, std::min<long long>(kAdjMaxParts, (kAdjTargetGroups + nb - 1) / nb));
  return (int)std::max<long long>(
      1, std::min<long long>(kFwdMaxParts, (kFwdTargetGroups + nb - 1) / nb));
}

Launch default_launch(bool adjoint, const Geo& g, int batch, bool ws = false) {
  Launch L;
  if (adjoint) {
    L.bx = 32;
    L.by = 16;
    L.bz = 16;
    const long long n32 = brick_count(g, 32, 16, 16, true) * batch;
    const long long n16 = brick_count(g, 16, 16, 16, true) * batch;
    if (ws ? n32 < 128 : n32 < 512) {
      L.bx = 16;
      L.by = 16;
      L.bz = (ws ? n16 >= 32 : n16 >= 200) ? 16 : 8;
    }
    L.parts = default_parts(true, g, batch, L.bx, L.by, L.bz, ws);
    return L;
  }
  if (brick_count(g, L.bx, L.by, L.bz, false) * batch * kFwdMaxParts < 256) {
    L.algo = DTV_ALGO_RAY;
    return L;
  }
  L.block = 512;
  L.parts = default_parts(false, g, batch, L.bx, L.by, L.bz);
  return L;
}

int resolve_launch(const dtv_launch* l, Launch& out, bool adjoint, const Geo& g, int batch,
                   bool ws = false) {
  out = default_launch(adjoint, g, batch, ws);
  if (!l) return DTV_OK;
  if (l->algorithm != DTV_ALGO_DEFAULT) {
    if (l->algorithm != DTV_ALGO_RAY && l->algorithm != DTV_ALGO_BRICK)
      return DTV_ERR_INVALID_ARGUMENT;
    out.algo = l->algorithm;
  }
  if (l->brick[0] || l->brick[1] || l->brick[2]) {
    for (int i = 0; i < 3; ++i)
      if (l->brick[i] < 2 || l->brick[i] > 64) return DTV_ERR_INVALID_ARGUMENT;
    if (fwd_tile_bytes(l->brick[0], l->brick[1], l->brick[2]) > kMaxTileBytes ||
        adj_tile_bytes(l->brick[0], l->brick[1], l->brick[2]) > kMaxTileBytes)
      return DTV_ERR_INVALID_ARGUMENT;
    out.bx = l->brick[0];
    out.by = l->brick[1];
    out.bz = l->brick[2];
    if (l->algorithm == DTV_ALGO_DEFAULT) out.algo = DTV_ALGO_BRICK;
  }
  if (l->block) {
    // A's brick kernel takes workgroups of up to 512 lanes, A^T's up to 256
    if (l->block < 64 || l->block > (adjoint ? 256 : 512) || l->block % 64)
      return DTV_ERR_INVALID_ARGUMENT;
    out.block = l->block;
  }
  if (l->view_parts) {
    const int maxp = adjoint ? (ws ? kAdjMaxParts : kAdjMaxAtomicParts) : kFwdMaxParts;
    if (l->view_parts < 1 || l->view_parts > maxp) return DTV_ERR_INVALID_ARGUMENT;
    out.parts = l->view_parts;
    out.parts_requested = true;
  } else {
    out.parts = default_parts(adjoint, g, batch, out.bx, out.by, out.bz, ws);
  }
  return DTV_OK;
}

int run_forward(const float* vol, float* sino, int32_t batch, const Geo& g, hipStream_t st,
                const Launch& L, void* fix_ws = nullptr);
int run_adjoint(const float* sino, float* vol, int32_t batch, const Geo& g, hipStream_t st,
                const Launch& L, float* ws, size_t ws_bytes);

int run_forward(const float* vol, float* sino, int32_t batch, const Geo& g, hipStream_t st) {
  return run_forward(vol, sino, batch, g, st, default_launch(false, g, batch));
}

// Bytes of the deterministic forward's workspace: a 256-byte header and one
// int64 per ray.
size_t forward_ws_bytes(int32_t batch, const Geo& g) {
  return 256 + (size_t)batch * g.n_views * g.nu * g.nv * sizeof(long long);
}

int run_forward(const float* vol, float* sino, int32_t batch, const Geo& g, hipStream_t st,
                const Launch& L, void* fix_ws) {
  if (L.algo == DTV_ALGO_BRICK) {
    int rc = check_atomic_target(sino);
    if (rc) return rc;
    if (fix_ws) {
      rc = check_atomic_target(fix_ws);
      if (rc) return rc;
      rc = run_zero(static_cast<float*>(fix_ws), forward_ws_bytes(batch, g) / sizeof(float), st);
      if (rc) return rc;
    }
    rc = run_zero(sino, (size_t)batch * g.n_views * g.nu * g.nv, st);
    if (rc) return rc;
    return launch_fwd_brick(vol, sino, batch, g, L.bx, L.by, L.bz, L.block, L.parts, st, fix_ws);
  }
  const dim3 block(kRayTileU, kRayTileV);
  const dim3 grid((g.nu + kRayTileU - 1) / kRayTileU, (g.nv + kRayTileV - 1) / kRayTileV,
                  g.n_views);
  hipLaunchKernelGGL(k_fwd_gather, grid, block, 0, st, vol, sino, g, (int)batch);
  return launch_status();
}

int run_adjoint(const float* sino, float* vol, int32_t batch, const Geo& g, hipStream_t st) {
  return run_adjoint(sino, vol, batch, g, st, default_launch(true, g, batch), nullptr, 0);
}

// Whether the workspace path partitions the samples over the bricks (its
// tile, with the upper-halo accumulators, fits the LDS limit).
bool adj_partitioned(const Launch& L) {
  return L.algo == DTV_ALGO_BRICK && L.bx * L.by * L.bz <= kAdjPerThread * L.block &&
         adj_ptile_bytes(L.bx, L.by, L.bz) <= kMaxTileBytes;
}

// Bytes of the A^T workspace: partitioned, the partial volumes of parts 1..
// and every part's upper-halo sums; else each part's partial volumes (none
// for one part).
size_t adjoint_ws_bytes(int32_t batch, const Geo& g, const Launch& L) {
  if (L.algo != DTV_ALGO_BRICK) return 0;
  if (adj_partitioned(L)) return adj_part_ws_bytes(batch, g, L.bx, L.by, L.bz, L.parts);
  return L.parts > 1 ? (size_t)L.parts * batch * g.nx * g.ny * g.nz * sizeof(float) : 0;
}

int run_adjoint(const float* sino, float* vol, int32_t batch, const Geo& g, hipStream_t st,
                const Launch& L, float* ws, size_t ws_bytes) {
  if (ws && adj_partitioned(L)) {
    // each sample marched by one brick; halos and parts summed in a fixed order
    if (ws_bytes < adjoint_ws_bytes(batch, g, L)) return DTV_ERR_INVALID_ARGUMENT;
    if (g.ny > 65535) return DTV_ERR_TOO_LARGE;
    // one part stores, two add into zeros when the output takes the hardware
    // atomics; else the parts go through the workspace
    bool into = L.parts == 1;
    if (L.parts == 2 && check_atomic_target(vol) == DTV_OK) {
      const int rc = run_zero(vol, (size_t)batch * g.nx * g.ny * g.nz, st);
      if (rc) return rc;
      into = true;
    }
    const int rc =
        launch_adj_part(sino, vol, ws, batch, g, L.bx, L.by, L.bz, L.block, L.parts, into, st);
    return rc == 0 ? DTV_OK : (rc == 1 ? DTV_ERR_INVALID_ARGUMENT : DTV_ERR_HIP);
  }
  if (L.algo == DTV_ALGO_BRICK && ws && L.parts > 1) {
    // every part stores its partial volumes; a second kernel adds them in order
    if (ws_bytes < adjoint_ws_bytes(batch, g, L)) return DTV_ERR_INVALID_ARGUMENT;
    return launch_adj_brick(sino, vol, ws, batch, g, L.bx, L.by, L.bz, L.block, L.parts, st);
  }
  if (L.algo == DTV_ALGO_BRICK) {
    int parts = L.parts;
    if (parts > kAdjMaxAtomicParts) return DTV_ERR_INVALID_ARGUMENT;  // needs a workspace
    if (parts > 1) {
      // the parts add into the output with FP32 global atomics
      int rc = check_atomic_target(vol);
      if (rc) {
        if (!L.parts_requested) {
          parts = 1;
        } else {
          return rc;
        }
      } else {
        rc = run_zero(vol, (size_t)batch * g.nx * g.ny * g.nz, st);
        if (rc) return rc;
      }
    }
    return launch_adj_brick(sino, vol, nullptr, batch, g, L.bx, L.by, L.bz, L.block, parts, st);
  }
  int rc = check_atomic_target(vol);
  if (rc) return rc;
  rc = run_zero(vol, (size_t)batch * g.nx * g.ny * g.nz, st);
  if (rc) return rc;
  const dim3 block(kRayTileU, kRayTileV);
  const dim3 grid((g.nu + kRayTileU - 1) / kRayTileU, (g.nv + kRayTileV - 1) / kRayTileV,
                  g.n_views);
  hipLaunchKernelGGL(k_adj_scatter, grid, block, 0, st, sino, vol, g, (int)batch);
  return launch_status();
}

}  // namespace

extern "C" {

int dtv_standard_forward(const float* volume, float* sino, int32_t batch, const double* angles,
                         int32_t n_angles, const int32_t volume_size[3],
                         const int32_t detector_size[2], double sid, double sdd,
                         int32_t device, void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_angles);
  if (rc) return rc;
  if (!volume || !sino || !angles) return DTV_ERR_NULL_POINTER;
  if (!(sid > 0.0) || !(sdd > sid)) return DTV_ERR_INVALID_ARGUMENT;
  DeviceGuard dg(device);
  if (dg.status) return dg.status;
  Geo g = make_geo(volume_size, detector_size, n_angles);
  scheme_standard(g, sid, sdd);
  g.kind = kStandard;
  g.angles = angles;
  return run_forward(volume, sino, batch, g, (hipStream_t)stream);
}

int dtv_standard_adjoint(const float* sino, float* volume, int32_t batch, const double* angles,
                         int32_t n_angles, const int32_t volume_size[3],
                         const int32_t detector_size[2], double sid, double sdd,
                         int32_t device, void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_angles);
  if (rc) return rc;
  if (!volume || !sino || !angles) return DTV_ERR_NULL_POINTER;
  if (!(sid > 0.0) || !(sdd > sid)) return DTV_ERR_INVALID_ARGUMENT;
  DeviceGuard dg(device);
  if (dg.status) return dg.status;
  Geo g = make_geo(volume_size, detector_size, n_angles);
  scheme_standard(g, sid, sdd);
  g.kind = kStandard;
  g.angles = angles;
  return run_adjoint(sino, volume, batch, g, (hipStream_t)stream);
}

int dtv_standard_backward(const float* sino, float* volume, int32_t batch, const double* angles,
                          int32_t n_angles, const int32_t volume_size[3],
                          const int32_t detector_size[2], double sid, double sdd,
                          int32_t device, void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_angles);
  if (rc) return rc;
  if (!volume || !sino || !angles) return DTV_ERR_NULL_POINTER;
  if (!(sid > 0.0) || !(sdd > sid)) return DTV_ERR_INVALID_ARGUMENT;
  DeviceGuard dg(device);
  if (dg.status) return dg.status;
  Geo g = make_geo(volume_size, detector_size, n_angles);
  scheme_standard(g, sid, sdd);
  g.kind = kStandard;
  g.angles = angles;
  const dim3 grid((g.nx + kBpTileX - 1) / kBpTileX, g.ny, (g.nz + kBpZ - 1) / kBpZ);
  for (int b = 0; b < batch; ++b) {
    hipLaunchKernelGGL(k_bp_standard, grid, dim3(kBpTileX), 0, (hipStream_t)stream, sino,
                       volume, g, b);
    rc = launch_status();
    if (rc) return rc;
  }
  return DTV_OK;
}

int dtv_beijing_forward(const float* volume, float* sino, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const float* project_vector, int32_t n_views, int32_t device,
                        void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!volume || !sino || !project_vector) return DTV_ERR_NULL_POINTER;
  DeviceGuard dg(device);
  if (dg.status) return dg.status;
  Geo g = make_geo(volume_size, detector_size, n_views);
  scheme_beijing(g);
  g.kind = kBeijing;
  g.pvec = project_vector;
  return run_forward(volume, sino, batch, g, (hipStream_t)stream);
}

int dtv_beijing_adjoint(const float* sino, float* volume, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const float* project_vector, int32_t n_views, int32_t device,
                        void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!volume || !sino || !project_vector) return DTV_ERR_NULL_POINTER;
  DeviceGuard dg(device);
  if (dg.status) return dg.status;
  Geo g = make_geo(volume_size, detector_size, n_views);
  scheme_beijing(g);
  g.kind = kBeijing;
  g.pvec = project_vector;
  return run_adjoint(sino, volume, batch, g, (hipStream_t)stream);
}

int dtv_beijing_backward(const float* sino, float* volume, int32_t batch,
                         const int32_t volume_size[3], const int32_t detector_size[2],
                         const float* project_vector, int32_t n_views, int32_t device,
                         void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!volume || !sino || !project_vector) return DTV_ERR_NULL_POINTER;
  if (n_views % 21 != 0) return DTV_ERR_VIEW_COUNT;
  DeviceGuard dg(device);
  if (dg.status) return dg.status;
  Geo g = make_geo(volume_size, detector_size, n_views);
  scheme_beijing(g);
  g.kind = kBeijing;
  g.pvec = project_vector;
  const dim3 grid((g.nx + kBpTileX - 1) / kBpTileX, g.ny, (g.nz + kBpZ - 1) / kBpZ);
  for (int b = 0; b < batch; ++b) {
    hipLaunchKernelGGL(k_bp_beijing, grid, dim3(kBpTileX), 0, (hipStream_t)stream, sino,
                       volume, g, b);
    rc = launch_status();
    if (rc) return rc;
  }
  return DTV_OK;
}

static int fill_table_geo(Geo& g, const int32_t* vs, const int32_t* ds, const double* views,
                          int32_t n_views, const dtv_scheme* sc) {
  if (!views || !sc) return DTV_ERR_NULL_POINTER;
  if (sc->axis_mode != kAxisXY && sc->axis_mode != kAxisXYZ) return DTV_ERR_INVALID_ARGUMENT;
  g = make_geo(vs, ds, n_views);
  g.axis_mode = sc->axis_mode;
  for (int c = 0; c < 3; ++c) {
    g.lo[c] = sc->box_lo[c];
    g.hi[c] = sc->box_hi[c];
  }
  g.margin = sc->margin;
  g.scale = sc->scale;
  g.kind = kTable;
  g.table = views;
  return DTV_OK;
}

int dtv_cone_forward(const float* volume, float* sino, int32_t batch,
                     const int32_t volume_size[3], const int32_t detector_size[2],
                     const double* views, int32_t n_views, const dtv_scheme* scheme,
                     void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!volume || !sino) return DTV_ERR_NULL_POINTER;
  Geo g;
  rc = fill_table_geo(g, volume_size, detector_size, views, n_views, scheme);
  if (rc) return rc;
  return run_forward(volume, sino, batch, g, (hipStream_t)stream);
}

int dtv_cone_adjoint(const float* sino, float* volume, int32_t batch,
                     const int32_t volume_size[3], const int32_t detector_size[2],
                     const double* views, int32_t n_views, const dtv_scheme* scheme,
                     void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!volume || !sino) return DTV_ERR_NULL_POINTER;
  Geo g;
  rc = fill_table_geo(g, volume_size, detector_size, views, n_views, scheme);
  if (rc) return rc;
  return run_adjoint(sino, volume, batch, g, (hipStream_t)stream);
}

int dtv_cone_forward_ex(const float* volume, float* sino, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const double* views, int32_t n_views, const dtv_scheme* scheme,
                        const dtv_launch* launch, void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!volume || !sino) return DTV_ERR_NULL_POINTER;
  Geo g;
  rc = fill_table_geo(g, volume_size, detector_size, views, n_views, scheme);
  if (rc) return rc;
  Launch L;
  rc = resolve_launch(launch, L, false, g, batch);
  if (rc) return rc;
  return run_forward(volume, sino, batch, g, (hipStream_t)stream, L);
}

int dtv_cone_forward_ws(const float* volume, float* sino, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const double* views, int32_t n_views, const dtv_scheme* scheme,
                        const dtv_launch* launch, void* workspace, size_t workspace_bytes,
                        void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!volume || !sino) return DTV_ERR_NULL_POINTER;
  Geo g;
  rc = fill_table_geo(g, volume_size, detector_size, views, n_views, scheme);
  if (rc) return rc;
  Launch L;
  rc = resolve_launch(launch, L, false, g, batch);
  if (rc) return rc;
  if (L.algo != DTV_ALGO_BRICK)  // the ray-driven gather is deterministic already
    return run_forward(volume, sino, batch, g, (hipStream_t)stream, L);
  if (!workspace) return DTV_ERR_NULL_POINTER;
  if (workspace_bytes < forward_ws_bytes(batch, g) || (uintptr_t)workspace % 256)
    return DTV_ERR_INVALID_ARGUMENT;
  return run_forward(volume, sino, batch, g, (hipStream_t)stream, L, workspace);
}

int dtv_cone_forward_plan(int32_t batch, const int32_t volume_size[3],
                          const int32_t detector_size[2], int32_t n_views,
                          const dtv_launch* launch, dtv_launch* plan, size_t* workspace_bytes) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!plan || !workspace_bytes) return DTV_ERR_NULL_POINTER;
  Geo g = make_geo(volume_size, detector_size, n_views);
  Launch L;
  rc = resolve_launch(launch, L, false, g, batch);
  if (rc) return rc;
  *plan = dtv_launch{};
  plan->algorithm = L.algo;
  plan->brick[0] = L.bx;
  plan->brick[1] = L.by;
  plan->brick[2] = L.bz;
  plan->block = L.block;
  plan->view_parts = L.algo == DTV_ALGO_BRICK ? L.parts : 1;
  *workspace_bytes = L.algo == DTV_ALGO_BRICK ? forward_ws_bytes(batch, g) : 0;
  return DTV_OK;
}

int dtv_cone_adjoint_ex(const float* sino, float* volume, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const double* views, int32_t n_views, const dtv_scheme* scheme,
                        const dtv_launch* launch, void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!volume || !sino) return DTV_ERR_NULL_POINTER;
  Geo g;
  rc = fill_table_geo(g, volume_size, detector_size, views, n_views, scheme);
  if (rc) return rc;
  Launch L;
  rc = resolve_launch(launch, L, true, g, batch);
  if (rc) return rc;
  return run_adjoint(sino, volume, batch, g, (hipStream_t)stream, L, nullptr, 0);
}

int dtv_cone_adjoint_ws(const float* sino, float* volume, int32_t batch,
                        const int32_t volume_size[3], const int32_t detector_size[2],
                        const double* views, int32_t n_views, const dtv_scheme* scheme,
                        const dtv_launch* launch, float* workspace, size_t workspace_bytes,
                        void* stream) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!volume || !sino) return DTV_ERR_NULL_POINTER;
  Geo g;
  rc = fill_table_geo(g, volume_size, detector_size, views, n_views, scheme);
  if (rc) return rc;
  Launch L;
  rc = resolve_launch(launch, L, true, g, batch, true);
  if (rc) return rc;
  // a NULL workspace is accepted for one part (the unpartitioned kernel)
  if (!workspace && L.algo == DTV_ALGO_BRICK && L.parts > 1) return DTV_ERR_NULL_POINTER;
  if (workspace && workspace_bytes < adjoint_ws_bytes(batch, g, L))
    return DTV_ERR_INVALID_ARGUMENT;
  return run_adjoint(sino, volume, batch, g, (hipStream_t)stream, L, workspace, workspace_bytes);
}

int dtv_cone_adjoint_plan(int32_t batch, const int32_t volume_size[3],
                          const int32_t detector_size[2], int32_t n_views,
                          const dtv_launch* launch, dtv_launch* plan, size_t* workspace_bytes) {
  int rc = check_sizes(batch, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!plan || !workspace_bytes) return DTV_ERR_NULL_POINTER;
  Geo g = make_geo(volume_size, detector_size, n_views);
  Launch L;
  rc = resolve_launch(launch, L, true, g, batch, true);
  if (rc) return rc;
  *plan = dtv_launch{};
  plan->algorithm = L.algo;
  plan->brick[0] = L.bx;
  plan->brick[1] = L.by;
  plan->brick[2] = L.bz;
  plan->block = L.block;
  plan->view_parts = L.algo == DTV_ALGO_BRICK ? L.parts : 1;
  *workspace_bytes = adjoint_ws_bytes(batch, g, L);
  return DTV_OK;
}

static void export_scheme(const Geo& g, dtv_scheme* out) {
  out->axis_mode = g.axis_mode;
  out->reserved = 0;
  for (int c = 0; c < 3; ++c) {
    out->box_lo[c] = g.lo[c];
    out->box_hi[c] = g.hi[c];
  }
  out->margin = g.margin;
  out->scale = g.scale;
}

int dtv_views_standard(const double* angles, int32_t n_angles, const int32_t volume_size[3],
                       const int32_t detector_size[2], double sid, double sdd,
                       double* views_out, dtv_scheme* scheme_out) {
  int rc = check_sizes(1, volume_size, detector_size, n_angles);
  if (rc) return rc;
  if (!angles || !views_out || !scheme_out) return DTV_ERR_NULL_POINTER;
  if (!(sid > 0.0) || !(sdd > sid)) return DTV_ERR_INVALID_ARGUMENT;
  Geo g = make_geo(volume_size, detector_size, n_angles);
  scheme_standard(g, sid, sdd);
  for (int a = 0; a < n_angles; ++a) {
    View v;
    view_standard(angles[a], sid, sdd, g.nx, g.ny, g.nz, g.nu, g.nv, v);
    double* o = views_out + 12 * (size_t)a;
    for (int c = 0; c < 3; ++c) {
      o[c] = v.S[c];
      o[3 + c] = v.D0[c];
      o[6 + c] = v.U[c];
      o[9 + c] = v.V[c];
    }
  }
  export_scheme(g, scheme_out);
  return DTV_OK;
}

int dtv_views_beijing(const float* project_vector, int32_t n_views, const int32_t volume_size[3],
                      const int32_t detector_size[2], double* views_out,
                      dtv_scheme* scheme_out) {
  int rc = check_sizes(1, volume_size, detector_size, n_views);
  if (rc) return rc;
  if (!project_vector || !views_out || !scheme_out) return DTV_ERR_NULL_POINTER;
  Geo g = make_geo(volume_size, detector_size, n_views);
  scheme_beijing(g);
  for (int a = 0; a < n_views; ++a) {
    View v;
    view_beijing(project_vector + 12 * (size_t)a, g.nx, g.ny, g.nz, g.nu, g.nv, v);
    double* o = views_out + 12 * (size_t)a;
    for (int c = 0; c < 3; ++c) {
      o[c] = v.S[c];
      o[3 + c] = v.D0[c];
      o[6 + c] = v.U[c];
      o[9 + c] = v.V[c];
    }
  }
  export_scheme(g, scheme_out);
  return DTV_OK;
}

const char* dtv_strerror(int code) {
  switch (code) {
    case DTV_OK: return "success";
    case DTV_ERR_INVALID_ARGUMENT: return "invalid argument (sizes, batch, view count or geometry)";
    case DTV_ERR_NULL_POINTER: return "null pointer argument";
    case DTV_ERR_TOO_LARGE: return "problem too large (32-bit per-volume indexing, the kernel's LDS budget or a fixed scratch capacity)";
    case DTV_ERR_HIP: return "HIP runtime error (launch or device selection failed)";
    case DTV_ERR_VIEW_COUNT: return "view count must be a multiple of 21 for the Beijing backprojector";
    default: return "unknown dtvproj error";
  }
}

// The build passes the digest of the sources it compiled (the loader compares
// it with the sources beside the library, so a stale build is refused).
#ifndef DTV_SOURCE_DIGEST
#define DTV_SOURCE_DIGEST unknown
#endif
#define DTV_STR2(x) #x
#define DTV_STR(x) DTV_STR2(x)

const char* dtv_version(void) { return "dtvproj 0.2.0 gfx950 src:" DTV_STR(DTV_SOURCE_DIGEST); }

}  // extern "C"

namespace dtv {

// Outputs that kernels accumulate into with hardware FP32 global atomics (A
// on bricks, the ray-driven A^T, the multi-instance A) must be coarse-grained
// device memory, as hipMalloc returns: on managed, host or fine-grained
// memory the hardware atomic add is not performed and sums come out wrong
// without any error.  Anything else, or a pointer HIP does not know, is
// rejected.
int check_atomic_target(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return DTV_ERR_INVALID_ARGUMENT;
  }
  // hipExtMallocWithFlags: 1 = fine-grained, 2 = signal memory, 3 = uncached
  if (a.type != hipMemoryTypeDevice || a.isManaged || (a.allocationFlags & 3u) != 0)
    return DTV_ERR_INVALID_ARGUMENT;
  return DTV_OK;
}

}  // namespace dtv
