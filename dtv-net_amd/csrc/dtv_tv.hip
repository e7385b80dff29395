// Fused smoothed-TV gradient step with non-negativity and momentum (gfx950).
//
// Per voxel c (flat C-order index i, x fastest):
//   d_a(i)  = f[i + s_a] - f[i]  (mode-dependent validity, below)
//   n(i)    = sqrt(d_x(i)^2 + d_y(i)^2 + d_z(i)^2 + eps^2)
//   gtv(i)  = sum_a [ prev_a(i) ? d_a(i - s_a) / n(i - s_a) : 0 ] - sum_a d_a(i) / n(i)
//   x       = y - step * (g + lam * gtv);  x = max(x, 0) if nonneg
//   y'      = x + beta * (x - x_prev)
// i.e. the gradient of TV_eps(y) = sum_i n(i) is D^T (D y / n).  Modes:
//   neumann   (0): s = (1, X, X*Y); d_a = 0 on the far face of axis a
//   reference (1): the sparse matrices of model/PDHG/Utils/Gradients.py:8-47:
//                  s = (1, Y, X*Y); D_x has no +1 where (i+1) % X == 0,
//                  D_y/D_z stop N - s before the end; rows without +1 give
//                  d = -f[i].
// One HBM pass: reads y (19-point stencil, cached), g, x_prev; writes x, y'.
// Uneven instances share one launch: each owns a contiguous range of
// workgroups of a one-dimensional grid, found by a binary search over the
// ranges (kernel arguments), so no workgroup falls outside its volume.
// Lanes take voxels of 64 x 4 x-y tiles (dtv_stencil.h), so neighbours and
// boundary tests come from coordinates, with no 64-bit division.
//
// The TV descent step of ASD-POCS (model/PDHG/ASDPOCS.py:37-43) uses the
// same stencil with another normaliser:
//   out = f - step * sum_a D_a^T (D_a f / (|D f| + eps)),
// the exact gradient step on sum_c [r_c - eps * log(1 + r_c / eps)],
// r_c = |D f|_c (the reference's eps is 1).
//
// The primal-dual (Chambolle-Pock) step of the directional-TV unroll runs
// as two launches after A and A^T: the dual box projection of
// p_a + sigma D_a fbar onto [-lam_a, lam_a] (the l1-ball dual step of
// model/PDHG/DTVCP.py:41-46 with a box in place of the ball), then the
// primal step f' = max(f - tau (A^T w + sum_a D_a^T p_a), 0) with the
// extrapolation fbar = 2 f' - f.  28 B per voxel each.
#include <hip/hip_runtime.h>

#include "dtv_stencil.h"
#include "dtvproj.h"

using namespace dtv;

namespace {

// q = D f / n at p: n = sqrt(|D f|^2 + eps^2) (kAsd false) or |D f| + eps
// (kAsd true)
template <bool kAsd>
__device__ __forceinline__ void qvec(const Vol3& v, const float* __restrict__ f, const Pos& p,
                                     float eps, float q[3]) {
  const float dx = diff(v, f, p, 0), dy = diff(v, f, p, 1), dz = diff(v, f, p, 2);
  const float ss = dx * dx + dy * dy + dz * dz;
  const float r = kAsd ? 1.f / (sqrtf(ss) + eps) : rsqrtf(ss + eps * eps);
  q[0] = dx * r;
  q[1] = dy * r;
  q[2] = dz * r;
}

// sum_a D_a^T q_a at p
template <bool kAsd>
__device__ __forceinline__ float tv_grad_at(const Vol3& v, const float* __restrict__ f,
                                            const Pos& p, float eps) {
  float q[3];
  qvec<kAsd>(v, f, p, eps, q);
  float gtv = -(q[0] + q[1] + q[2]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (has_prev(v, p, a)) {
      float qp[3];
      qvec<kAsd>(v, f, back(v, p, a), eps, qp);
      gtv += qp[a];
    }
  }
  return gtv;
}

__device__ __forceinline__ void tv_point(const dtv_tv_item& it, const Vol3& v, int64_t off,
                                         const Pos& p) {
  const float* y = it.y + off;
  const float gtv = tv_grad_at<false>(v, y, p, it.eps);
  // read every input at i before writing: x_prev may share storage with an
  // output (y may not: neighbours read it)
  const int64_t i = off + p.i;
  const float yv = y[p.i];
  const float gv = it.g[i];
  const float xp = it.xprev[i];
  float x = yv - it.step * (gv + it.lam * gtv);
  if (it.nonneg) x = fmaxf(x, 0.f);
  it.xnew[i] = x;
  it.ynew[i] = x + it.beta * (x - xp);
}

// one instance, blockIdx.z = batch item (volumes of equal size, contiguous)
__global__ __launch_bounds__(256) void k_tv_fista_one(dtv_tv_item it, Vol3 v) {
  Pos p;
  if (!vol3_pos(v, p)) return;
  tv_point(it, v, (int64_t)blockIdx.z * v.n, p);
}

// Uneven instances in one launch: instance k owns the blocks
// [off[k], off[k+1]) of a one-dimensional grid, (x-y tiles) x z of its own
// volume, so no workgroup is launched outside an instance.
constexpr int kMaxTvItems = 256;
struct TvOffsets {
  int n;
  int off[kMaxTvItems + 1];
};

__global__ __launch_bounds__(256) void k_tv_fista_multi(const dtv_tv_item* __restrict__ items,
                                                        int mode, TvOffsets offs) {
  const int b = (int)blockIdx.x;
  int lo = 0, hi = offs.n;  // last k with off[k] <= b
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offs.off[mid] <= b) lo = mid;
    else hi = mid;
  }
  const dtv_tv_item it = items[lo];
  const Vol3 v = make_vol3(it.nx, it.ny, it.nz, mode);
  const int tiles = v.tiles_x * ((v.ny + kTileY - 1) / kTileY);
  const int t = b - offs.off[lo];
  const int z = t / tiles;
  const int tile = t - z * tiles;
  const int ty = tile / v.tiles_x, tx = tile - ty * v.tiles_x;
  Pos p;
  p.x = tx * kTileX + (int)(threadIdx.x & (kTileX - 1));
  p.y = ty * kTileY + (int)(threadIdx.x / kTileX);
  p.z = z;
  p.i = ((int64_t)p.z * v.ny + p.y) * v.nx + p.x;
  if (p.x >= v.nx || p.y >= v.ny) return;
  tv_point(it, v, 0, p);
}

// out = f - step * TV gradient (ASD normaliser), blockIdx.z = batch item
__global__ __launch_bounds__(256) void k_tv_descent(const float* __restrict__ f,
                                                    float* __restrict__ out, Vol3 v,
                                                    float step, float eps) {
  Pos p;
  if (!vol3_pos(v, p)) return;
  const int64_t off = (int64_t)blockIdx.z * v.n;
  const float* fb = f + off;
  out[off + p.i] = fb[p.i] - step * tv_grad_at<true>(v, fb, p, eps);
}

// ---- z-marching form of the TV gradient ---------------------------------------
//
// The gradient at i needs q = D f / n at i and at its three back neighbours
// i - s_a.  Evaluated per output voxel that is four q's (and 13 loads of f);
// here every q is evaluated once: a workgroup of 64 x 8 lanes covers the x-y
// region [x0 - 1, x0 + 63) x [y0 - 1, y0 + 7) and marches z through a chunk
// of slices.  Per slice each lane evaluates q at its point, q_x and q_y go to
// LDS for the lanes at x + 1 and y + 1, q_z stays in a register for the next
// slice, and the 63 x 7 lanes off the region's low edges produce their voxel.
// That needs i - s_1 to be the voxel at y - 1: always in Neumann mode; in the
// reference mode (s_1 = Y) when X == Y, where at y = 0 the back neighbour is
// (x, Y - 1, z - 1), evaluated directly.  Reference mode with X != Y takes
// the per-voxel kernels above.
constexpr int kMX = 64, kMY = 8, kMZ = 8;  // march region and z chunk (8 slices: more workgroups in flight; 32, 16, 6 and 4 measured slower)

__host__ __device__ inline bool march_ok(const Vol3& v) {
  return (v.mode == 0 || v.nx == v.ny) && v.n < (1ll << 31);
}
__host__ __device__ inline int march_tiles_x(const Vol3& v) { return (v.nx + kMX - 2) / (kMX - 1); }
__host__ __device__ inline int march_tiles_y(const Vol3& v) { return (v.ny + kMY - 2) / (kMY - 1); }
__host__ __device__ inline int march_chunks(const Vol3& v) { return (v.nz + kMZ - 1) / kMZ; }
__host__ __device__ inline int march_blocks(const Vol3& v) {
  return march_tiles_x(v) * march_tiles_y(v) * march_chunks(v);
}

// q = D f / n at a point from its value fc and those of its three forward
// neighbours (read only where the point has them).
template <bool kAsd>
__device__ __forceinline__ void qvec_vals(const Vol3& v, bool hx, bool hy, bool hz, float fc,
                                          float fx, float fy, float fz, float eps, float q[3]) {
  const float edge = v.mode == 0 ? 0.f : -fc;
  const float dx = hx ? fx - fc : edge, dy = hy ? fy - fc : edge, dz = hz ? fz - fc : edge;
  const float ss = dx * dx + dy * dy + dz * dz;
  const float r = kAsd ? 1.f / (sqrtf(ss) + eps) : rsqrtf(ss + eps * eps);
  q[0] = dx * r;
  q[1] = dy * r;
  q[2] = dz * r;
}

// Marches block b of volume f (march_blocks(v) blocks per volume).  Per
// voxel the lane produces, out.load(i) fetches what the update reads besides
// f (issued one slice ahead) and out.store(i, fc, gtv, pre) writes it; gtv =
// sum_a D_a^T q_a and fc = f(i).  Each lane holds f at its point of the
// current slice and of the next one (loaded a slice ahead): the x neighbour
// is the next lane's value (a wave is one row of the region), the y
// neighbour the next row's value, which every lane copies into LDS a slice
// ahead, so per slice a lane loads one value of f (lane 63 and the last row
// load their neighbour beyond the region, and in the reference mode the
// last row of the volume its neighbour (x, 0, z + 1)).  Offsets are 32-bit
// (volumes below 2^31 voxels) from the volume's base; q and f of two
// consecutive slices use alternate LDS buffers, so one barrier per slice
// orders them.
// Pointers read from a descriptor table in memory are generic (flat) to the
// compiler, and flat loads count against the LDS wait counter as well, so
// every LDS access in the march would also wait for the prefetched slice: the
// march and the FISTA output address global memory explicitly.
typedef __attribute__((address_space(1))) float gfloat;
__device__ __forceinline__ const gfloat* glob(const float* p) { return (const gfloat*)p; }
__device__ __forceinline__ gfloat* glob(float* p) { return (gfloat*)p; }

template <bool kAsd, typename Out>
__device__ __forceinline__ void tv_march(const Vol3& v, const float* __restrict__ f_, float eps,
                                         int b, Out& out) {
  const gfloat* __restrict__ f = glob(f_);
  __shared__ float sq[2][2][kMY][kMX];
  __shared__ float sf[2][kMY][kMX];
  const int txy = march_tiles_x(v) * march_tiles_y(v);
  const int chunk = b / txy, t = b - chunk * txy;
  const int ty = t / march_tiles_x(v), tx = t - ty * march_tiles_x(v);
  const int lx = threadIdx.x & (kMX - 1), ly = threadIdx.x / kMX;
  const int x = tx * (kMX - 1) - 1 + lx, y = ty * (kMY - 1) - 1 + ly;
  const bool in_xy = x >= 0 && x < v.nx && y >= 0 && y < v.ny;
  const bool out_lane = in_xy && lx >= 1 && ly >= 1;
  const int z0 = chunk * kMZ, z1 = min(v.nz, z0 + kMZ);
  const unsigned s1 = (unsigned)v.s1, s2 = (unsigned)v.s2, n = (unsigned)v.n;
  const unsigned row = (unsigned)(y * v.nx + x);
  const bool hx = x != v.nx - 1;
  const bool ref = v.mode != 0;
  // has_next on D_y: Neumann y != Y - 1, reference i < n - s_1
  auto hy_at = [&](unsigned i) { return ref ? i < n - s1 : y != v.ny - 1; };
  // where the y neighbour comes from: the next row's LDS copy, or memory
  const bool fy_lds = ly < kMY - 1 && y + 1 < v.ny;
  const bool fx_mem = lx == kMX - 1 && hx;
  float qzp = 0.f, fc = 0.f;
  if (in_xy) {
    const unsigned i0 = (unsigned)z0 * s2 + row;
    fc = f[i0];
    if (z0 > 0) {
      const unsigned ib = i0 - s2;
      float q[3];
      qvec_vals<kAsd>(v, hx, hy_at(ib), true, f[ib], hx ? f[ib + 1] : 0.f,
                      hy_at(ib) ? f[ib + s1] : 0.f, fc, eps, q);
      qzp = q[2];
    }
  }
  sf[z0 & 1][ly][lx] = fc;
  // what slice z needs from memory is loaded during slice z - 1 (software
  // pipelining): f one slice further up, the neighbours beyond the region
  float fxm = 0.f, fym = 0.f, fz = 0.f;
  typename Out::Pre pre{};
  auto fetch = [&](int z, float& gx, float& gy, float& gz, typename Out::Pre& gp) {
    const unsigned i = (unsigned)z * s2 + row;
    if (fx_mem) gx = f[i + 1];
    if (!fy_lds && hy_at(i)) gy = f[i + s1];
    if (z != v.nz - 1) gz = f[i + s2];
    if (out_lane) gp = out.load(i);
  };
  if (in_xy && z0 < z1) fetch(z0, fxm, fym, fz, pre);
  __syncthreads();  // the first slice's f in sf
  // fully unrolled, so the prefetched values change registers by renaming
  // rather than by copies (a copy waits for its load within the iteration)
#pragma unroll
  for (int k = 0; k < kMZ; ++k) {
    const int z = z0 + k;
    if (z >= z1) break;
    const unsigned i = (unsigned)z * s2 + row;
    const bool hy = hy_at(i), hz = z != v.nz - 1;
    float nfx = 0.f, nfy = 0.f, nfz = 0.f;
    typename Out::Pre npre{};
    if (in_xy && z + 1 < z1) fetch(z + 1, nfx, nfy, nfz, npre);
    const int buf = z & 1;
    const float right = __shfl_down(fc, 1, kMX);
    const float fx = fx_mem ? fxm : right;
    const float fy = fy_lds ? sf[buf][ly + 1][lx] : fym;
    float q[3] = {0.f, 0.f, 0.f};
    if (in_xy) qvec_vals<kAsd>(v, hx, hy, hz, fc, fx, fy, fz, eps, q);
    sq[buf][0][ly][lx] = q[0];
    sq[buf][1][ly][lx] = q[1];
    sf[buf ^ 1][ly][lx] = fz;  // read by the row below during slice z + 1
    __syncthreads();
    if (out_lane) {
      float gtv = -(q[0] + q[1] + q[2]);
      if (x > 0) gtv += sq[buf][0][ly][lx - 1];
      if (y > 0) {
        gtv += sq[buf][1][ly - 1][lx];
      } else if (ref && z > 0) {
        // reference D_y with X == Y: row i - Y is (x, Y - 1, z - 1)
        const unsigned ib = i - s1;
        const bool hyb = hy_at(ib);
        float qb[3];
        qvec_vals<kAsd>(v, hx, hyb, true, f[ib], hx ? f[ib + 1] : 0.f, hyb ? f[ib + s1] : 0.f,
                        f[ib + s2], eps, qb);
        gtv += qb[1];
      }
      if (z > 0) gtv += qzp;
      out.store(i, fc, gtv, pre);
    }
    qzp = q[2];
    fc = fz;
    fxm = nfx;
    fym = nfy;
    fz = nfz;
    pre = npre;
  }
}

// FISTA-TV update of one volume (base pointers of this instance / item)
struct FistaOut {
  struct Pre {
    float g, xp;
  };
  const float* g;
  const float* xprev;
  float* xnew;
  float* ynew;
  float step, lam, beta;
  int nonneg;
  __device__ __forceinline__ Pre load(unsigned i) const {
    return Pre{glob(g)[i], glob(xprev)[i]};
  }
  __device__ __forceinline__ void store(unsigned i, float yv, float gtv, const Pre& p) const {
    float x = yv - step * (p.g + lam * gtv);
    if (nonneg) x = fmaxf(x, 0.f);
    glob(xnew)[i] = x;
    glob(ynew)[i] = x + beta * (x - p.xp);
  }
};

__device__ __forceinline__ FistaOut fista_out(const dtv_tv_item& it, int64_t off) {
  return FistaOut{it.g + off, it.xprev + off, it.xnew + off, it.ynew + off, it.step, it.lam,
                  it.beta, it.nonneg};
}

// FISTA-TV step, one instance: blockIdx.x = march block, blockIdx.y = batch item
__global__ __launch_bounds__(512) void k_tv_fista_march(dtv_tv_item it, Vol3 v) {
  const int64_t off = (int64_t)blockIdx.y * v.n;
  FistaOut o = fista_out(it, off);
  tv_march<false>(v, it.y + off, it.eps, (int)blockIdx.x, o);
}

struct DescentOut {
  struct Pre {};
  float* out;
  float step;
  __device__ __forceinline__ Pre load(unsigned) const { return Pre{}; }
  __device__ __forceinline__ void store(unsigned i, float fv, float gtv, const Pre&) const {
    out[i] = fv - step * gtv;
  }
};

// ASD-POCS TV descent: blockIdx.x = march block, blockIdx.y = batch item
__global__ __launch_bounds__(512) void k_tv_descent_march(const float* __restrict__ f,
                                                          float* __restrict__ out, Vol3 v,
                                                          float step, float eps) {
  const int64_t off = (int64_t)blockIdx.y * v.n;
  DescentOut o{out + off, step};
  tv_march<true>(v, f + off, eps, (int)blockIdx.x, o);
}

// ---- z march, four voxels per lane ------------------------------------------------
//
// With X a multiple of 4 (and 16-byte aligned arrays) each lane of the march
// takes four consecutive x voxels as one float4: a 64 x 8 workgroup covers
// [x0, x0 + 256) x [y0 - 1, y0 + 7); the lanes of one wave are one row, so
// q_x of the voxel before a quad and f of the voxel after it come from the
// neighbouring lanes by shuffles (the first lane evaluates that q_x itself,
// the last lane loads that f), and only q_y goes through LDS.
__host__ __device__ inline int march4_tiles_x(const Vol3& v) { return (v.nx + 255) / 256; }
// taken when X is a multiple of 4 and the 256-wide tiles keep at least as
// many lanes busy as the one-voxel march's 63-wide tiles (up to 5%): a row
// of 124 voxels, say, would leave half of a wave idle
__host__ __device__ inline bool march4_ok(const Vol3& v) {
  if (!march_ok(v) || v.nx % 4 != 0) return false;
  const float u4 = (float)v.nx / (256.f * (float)march4_tiles_x(v));
  const float u1 = (float)v.nx / (63.f * (float)((v.nx + 62) / 63));
  return u4 >= 0.95f * u1;
}
__host__ __device__ inline int march4_blocks(const Vol3& v) {
  return march4_tiles_x(v) * march_tiles_y(v) * march_chunks(v);
}

__device__ __forceinline__ float4 ld4(const float* __restrict__ f, unsigned i) {
  return *reinterpret_cast<const float4*>(f + i);
}
__device__ __forceinline__ float el(const float4& a, int e) {
  return e == 0 ? a.x : (e == 1 ? a.y : (e == 2 ? a.z : a.w));
}

// q of a quad's four voxels from its values c, the value after it (xa, used
// when the last voxel has an x neighbour: hx3), the quad one step along y
// (yn, when hy) and along z (zn, when hz)
template <bool kAsd>
__device__ __forceinline__ void qvec4(bool ref, bool hx3, bool hy, bool hz, const float4& c,
                                      float xa, const float4& yn, const float4& zn, float eps,
                                      float qx[4], float qy[4], float qz[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const float ce = el(c, e);
    const float edge = ref ? -ce : 0.f;
    const float xn = e < 3 ? el(c, e + 1) : xa;
    const float dx = (e < 3 || hx3) ? xn - ce : edge;
    const float dy = hy ? el(yn, e) - ce : edge;
    const float dz = hz ? el(zn, e) - ce : edge;
    const float ss = dx * dx + dy * dy + dz * dz;
    const float r = kAsd ? 1.f / (sqrtf(ss) + eps) : rsqrtf(ss + eps * eps);
    qx[e] = dx * r;
    qy[e] = dy * r;
    qz[e] = dz * r;
  }
}

template <bool kAsd, typename Out>
__device__ __forceinline__ void tv_march4(const Vol3& v, const float* __restrict__ f, float eps,
                                          int b, Out& out) {
  __shared__ float4 sqy[2][kMY][kMX];
  const int tq = march4_tiles_x(v);
  const int txy = tq * march_tiles_y(v);
  const int chunk = b / txy, t = b - chunk * txy;
  const int ty = t / tq, tx = t - ty * tq;
  const int lx = threadIdx.x & (kMX - 1), ly = threadIdx.x / kMX;
  const int x = tx * 256 + 4 * lx, y = ty * (kMY - 1) - 1 + ly;
  const bool in_xy = x < v.nx && y >= 0 && y < v.ny;
  const bool out_lane = in_xy && ly >= 1;
  const int z0 = chunk * kMZ, z1 = min(v.nz, z0 + kMZ);
  const unsigned s1 = (unsigned)v.s1, s2 = (unsigned)v.s2, n = (unsigned)v.n;
  const unsigned row = in_xy ? (unsigned)(y * v.nx + x) : 0u;
  const bool hx3 = x + 4 < v.nx;
  const bool ref = v.mode != 0;
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  auto hy_at = [&](unsigned i) { return ref ? i < n - s1 : y != v.ny - 1; };
  float qzp[4] = {0.f, 0.f, 0.f, 0.f};
  float4 c = zero4;
  if (in_xy) {
    const unsigned i0 = (unsigned)z0 * s2 + row;
    c = ld4(f, i0);
    if (z0 > 0) {
      const unsigned ib = i0 - s2;
      const bool hyb = hy_at(ib);
      float qx[4], qy[4];
      qvec4<kAsd>(ref, hx3, hyb, true, ld4(f, ib), hx3 ? f[ib + 4] : 0.f,
                  hyb ? ld4(f, ib + s1) : zero4, c, eps, qx, qy, qzp);
    }
  }
  // the values of slice z are loaded during slice z - 1
  float4 yn = zero4, zn = zero4;
  float xa = 0.f;
  typename Out::Pre pre{};
  auto fetch = [&](int z, float4& gy, float4& gz, float& gx, typename Out::Pre& gp) {
    const unsigned i = (unsigned)z * s2 + row;
    if (hy_at(i)) gy = ld4(f, i + s1);
    if (z != v.nz - 1) gz = ld4(f, i + s2);
    if (lx == kMX - 1 && hx3) gx = f[i + 4];
    if (out_lane) gp = out.load(i);
  };
  if (in_xy && z0 < z1) fetch(z0, yn, zn, xa, pre);
#pragma unroll
  for (int k = 0; k < kMZ; ++k) {
    const int z = z0 + k;
    if (z >= z1) break;
    const unsigned i = (unsigned)z * s2 + row;
    const bool hy = in_xy && hy_at(i), hz = z != v.nz - 1;
    float4 nyn = zero4, nzn = zero4;
    float nxa = 0.f;
    typename Out::Pre npre{};
    if (in_xy && z + 1 < z1) fetch(z + 1, nyn, nzn, nxa, npre);
    // the value after the quad: the next lane's first (the last lane loaded it)
    const float after = __shfl_down(c.x, 1, kMX);
    float qx[4] = {0.f, 0.f, 0.f, 0.f}, qy[4] = {0.f, 0.f, 0.f, 0.f}, qz[4] = {0.f, 0.f, 0.f, 0.f};
    if (in_xy)
      qvec4<kAsd>(ref, hx3, hy, hz, c, lx == kMX - 1 ? xa : after, yn, zn, eps, qx, qy, qz);
    // q_x of the voxel before the quad: the previous lane's last, or for the
    // first lane of the row evaluated here
    float qx_before = __shfl_up(qx[3], 1, kMX);
    if (lx == 0 && out_lane && x > 0) {
      const unsigned ia = i - 1;
      const bool hya = hy_at(ia);
      const float fa = f[ia];
      const float edge = ref ? -fa : 0.f;
      const float dx = c.x - fa;
      const float dy = hya ? f[ia + s1] - fa : edge;
      const float dz = hz ? f[ia + s2] - fa : edge;
      const float ss = dx * dx + dy * dy + dz * dz;
      qx_before = dx * (kAsd ? 1.f / (sqrtf(ss) + eps) : rsqrtf(ss + eps * eps));
    }
    const int buf = z & 1;
    sqy[buf][ly][lx] = make_float4(qy[0], qy[1], qy[2], qy[3]);
    __syncthreads();
    if (out_lane) {
      float qyb[4] = {0.f, 0.f, 0.f, 0.f};
      bool has_yb = false;
      if (y > 0) {
        const float4 w = sqy[buf][ly - 1][lx];
        qyb[0] = w.x;
        qyb[1] = w.y;
        qyb[2] = w.z;
        qyb[3] = w.w;
        has_yb = true;
      } else if (ref && z > 0) {
        // reference D_y with X == Y: row i - Y is (x, Y - 1, z - 1)
        const unsigned ib = i - s1;
        float bx[4], bz[4];
        qvec4<kAsd>(ref, hx3, hy_at(ib), true, ld4(f, ib), hx3 ? f[ib + 4] : 0.f, c,
                    ld4(f, ib + s2), eps, bx, qyb, bz);
        has_yb = true;
      }
      float gtv[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float gv = -(qx[e] + qy[e] + qz[e]);
        if (e > 0) gv += qx[e - 1];
        else if (x > 0) gv += qx_before;
        if (has_yb) gv += qyb[e];
        if (z > 0) gv += qzp[e];
        gtv[e] = gv;
      }
      out.store4(i, c, gtv, pre);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) qzp[e] = qz[e];
    c = zn;
    yn = nyn;
    zn = nzn;
    xa = nxa;
    pre = npre;
  }
}

struct FistaOut4 {
  struct Pre {
    float4 g, xp;
  };
  const float* g;
  const float* xprev;
  float* xnew;
  float* ynew;
  float step, lam, beta;
  int nonneg;
  __device__ __forceinline__ Pre load(unsigned i) const { return Pre{ld4(g, i), ld4(xprev, i)}; }
  __device__ __forceinline__ void store4(unsigned i, const float4& yv, const float gtv[4],
                                         const Pre& p) const {
    float x[4], yy[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float xe = el(yv, e) - step * (el(p.g, e) + lam * gtv[e]);
      if (nonneg) xe = fmaxf(xe, 0.f);
      x[e] = xe;
      yy[e] = xe + beta * (xe - el(p.xp, e));
    }
    *reinterpret_cast<float4*>(xnew + i) = make_float4(x[0], x[1], x[2], x[3]);
    *reinterpret_cast<float4*>(ynew + i) = make_float4(yy[0], yy[1], yy[2], yy[3]);
  }
};

__device__ __forceinline__ FistaOut4 fista_out4(const dtv_tv_item& it, int64_t off) {
  return FistaOut4{it.g + off, it.xprev + off, it.xnew + off, it.ynew + off, it.step, it.lam,
                   it.beta, it.nonneg};
}

__global__ __launch_bounds__(512) void k_tv_fista_march4(dtv_tv_item it, Vol3 v) {
  const int64_t off = (int64_t)blockIdx.y * v.n;
  FistaOut4 o = fista_out4(it, off);
  tv_march4<false>(v, it.y + off, it.eps, (int)blockIdx.x, o);
}

// uneven instances: instance k owns march blocks [off[k], off[k+1]).  All
// take the one-voxel march: one kernel holding both march bodies carries the
// larger one's registers and LDS for every workgroup, which measured slower
// on the 16-instance workload than the one-voxel march alone.
__global__ __launch_bounds__(512) void k_tv_fista_march_multi(
    const dtv_tv_item* __restrict__ items, int mode, TvOffsets offs) {
  const int b = (int)blockIdx.x;
  int lo = 0, hi = offs.n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offs.off[mid] <= b) lo = mid;
    else hi = mid;
  }
  const dtv_tv_item it = items[lo];
  const Vol3 v = make_vol3(it.nx, it.ny, it.nz, mode);
  FistaOut o = fista_out(it, 0);
  tv_march<false>(v, it.y, it.eps, b - offs.off[lo], o);
}

struct DescentOut4 {
  struct Pre {};
  float* out;
  float step;
  __device__ __forceinline__ Pre load(unsigned) const { return Pre{}; }
  __device__ __forceinline__ void store4(unsigned i, const float4& fv, const float gtv[4],
                                         const Pre&) const {
    *reinterpret_cast<float4*>(out + i) = make_float4(fv.x - step * gtv[0], fv.y - step * gtv[1],
                                                      fv.z - step * gtv[2], fv.w - step * gtv[3]);
  }
};

__global__ __launch_bounds__(512) void k_tv_descent_march4(const float* __restrict__ f,
                                                           float* __restrict__ out, Vol3 v,
                                                           float step, float eps) {
  const int64_t off = (int64_t)blockIdx.y * v.n;
  DescentOut4 o{out + off, step};
  tv_march4<true>(v, f + off, eps, (int)blockIdx.x, o);
}

// one PDHG step of the directional-TV unroll, after A and A^T.  Dual half:
//   p_a <- clamp(p_a + sigma * d_a(fbar), -lam_a, lam_a)
// (p: three planes [3][batch][Z][Y][X], updated in place; each lane reads
// only its own p entries).
__global__ __launch_bounds__(256) void k_pdhg_dual(const float* __restrict__ fbar,
                                                   float* __restrict__ p, Vol3 v,
                                                   int64_t plane, float sigma, float l0,
                                                   float l1, float l2) {
  Pos q;
  if (!vol3_pos(v, q)) return;
  const int64_t off = (int64_t)blockIdx.z * v.n;
  const float* fb = fbar + off;
  const float lam[3] = {l0, l1, l2};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float* pa = p + a * plane + off;
    const float val = pa[q.i] + sigma * diff(v, fb, q, a);
    pa[q.i] = fminf(fmaxf(val, -lam[a]), lam[a]);
  }
}

// Primal half: g = atw + sum_a D_a^T p_a;  f_new = max(f - tau g, 0);
// fbar <- 2 f_new - f;  f <- f_new.
__global__ __launch_bounds__(256) void k_pdhg_primal(float* __restrict__ f,
                                                     float* __restrict__ fbar,
                                                     const float* __restrict__ atw,
                                                     const float* __restrict__ p, Vol3 v,
                                                     int64_t plane, float tau) {
  Pos q;
  if (!vol3_pos(v, q)) return;
  const int64_t off = (int64_t)blockIdx.z * v.n;
  const int64_t i = off + q.i;
  float g = atw[i];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float* pa = p + a * plane + off;
    float dt = has_prev(v, q, a) ? pa[q.i - stride(v, a)] : 0.f;
    if (has_self(v, q, a)) dt -= pa[q.i];
    g += dt;
  }
  const float fo = f[i];
  const float fn = fmaxf(fo - tau * g, 0.f);
  f[i] = fn;
  fbar[i] = 2.f * fn - fo;
}

// ---- four voxels per lane --------------------------------------------------------
//
// When X (and in reference mode Y) is a multiple of 4, rows and the D_y / D_z
// strides keep 16-byte alignment, so each lane takes four consecutive x
// voxels with float4 loads and stores: workgroups of 64 x 4 lanes cover 256
// x-voxels of 4 rows, blockIdx.y = z, blockIdx.z = batch item.  The boundary
// tests on D_y and D_z hold for all four voxels at once (their strides and
// the array ends are multiples of 4); D_x needs the next quad's first value.
struct Quad {
  int x, y, z;  // x = first of the four
  unsigned i;   // flat index of (x, y, z) within the volume
};

__host__ __device__ inline bool quad_ok(const Vol3& v) {
  return v.nx % 4 == 0 && (v.mode == 0 || v.ny % 4 == 0) && v.n < (1ll << 31);
}
__host__ __device__ inline int quad_tiles_x(const Vol3& v) { return (v.nx + 255) / 256; }
inline dim3 quad_grid(const Vol3& v, int batch) {
  return dim3((unsigned)(quad_tiles_x(v) * ((v.ny + 3) / 4)), (unsigned)v.nz, (unsigned)batch);
}
__device__ __forceinline__ bool quad_pos(const Vol3& v, Quad& q) {
  const int tq = quad_tiles_x(v);
  const int ty = (int)blockIdx.x / tq, tx = (int)blockIdx.x - ty * tq;
  q.x = tx * 256 + 4 * (int)(threadIdx.x & 63);
  q.y = ty * 4 + (int)(threadIdx.x >> 6);
  q.z = (int)blockIdx.y;
  q.i = (unsigned)((q.z * v.ny + q.y) * v.nx + q.x);
  return q.x < v.nx && q.y < v.ny;
}
// has_next on D_y / D_z (the same for the quad's four voxels)
__device__ __forceinline__ bool quad_next(const Vol3& v, const Quad& q, int a) {
  if (v.mode == 0) return a == 1 ? q.y != v.ny - 1 : q.z != v.nz - 1;
  return (int64_t)q.i < v.n - stride(v, a);
}
__device__ __forceinline__ bool quad_prev(const Vol3& v, const Quad& q, int a) {
  if (v.mode == 0) return a == 1 ? q.y != 0 : q.z != 0;
  return (int64_t)q.i >= stride(v, a);
}

__global__ __launch_bounds__(256) void k_pdhg_dual4(const float* __restrict__ fbar,
                                                    float* __restrict__ p, Vol3 v,
                                                    int64_t plane, float sigma, float l0,
                                                    float l1, float l2) {
  Quad q;
  if (!quad_pos(v, q)) return;
  const int64_t off = (int64_t)blockIdx.z * v.n;
  const float* fb = fbar + off;
  const float4 c = *reinterpret_cast<const float4*>(fb + q.i);
  const float edge0 = v.mode == 0 ? 0.f : -c.w;
  // D_x: within the quad, and from the next quad's first value at the end
  const bool last = q.x + 4 == v.nx;
  const float4 dx = make_float4(c.y - c.x, c.z - c.y, c.w - c.z, last ? edge0 : fb[q.i + 4] - c.w);
  float4 dy, dz;
  const float4 cn = make_float4(-c.x, -c.y, -c.z, -c.w);
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
  if (quad_next(v, q, 1)) {
    const float4 ny = *reinterpret_cast<const float4*>(fb + q.i + (unsigned)v.s1);
    dy = make_float4(ny.x - c.x, ny.y - c.y, ny.z - c.z, ny.w - c.w);
  } else {
    dy = v.mode == 0 ? zero : cn;
  }
  if (quad_next(v, q, 2)) {
    const float4 nz = *reinterpret_cast<const float4*>(fb + q.i + (unsigned)v.s2);
    dz = make_float4(nz.x - c.x, nz.y - c.y, nz.z - c.z, nz.w - c.w);
  } else {
    dz = v.mode == 0 ? zero : cn;
  }
  const float4 d[3] = {dx, dy, dz};
  const float lam[3] = {l0, l1, l2};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float4* pa = reinterpret_cast<float4*>(p + a * plane + off + q.i);
    float4 w = *pa;
    const float l = lam[a];
    w.x = fminf(fmaxf(w.x + sigma * d[a].x, -l), l);
    w.y = fminf(fmaxf(w.y + sigma * d[a].y, -l), l);
    w.z = fminf(fmaxf(w.z + sigma * d[a].z, -l), l);
    w.w = fminf(fmaxf(w.w + sigma * d[a].w, -l), l);
    *pa = w;
  }
}

__global__ __launch_bounds__(256) void k_pdhg_primal4(float* __restrict__ f,
                                                      float* __restrict__ fbar,
                                                      const float* __restrict__ atw,
                                                      const float* __restrict__ p, Vol3 v,
                                                      int64_t plane, float tau) {
  Quad q;
  if (!quad_pos(v, q)) return;
  const int64_t off = (int64_t)blockIdx.z * v.n;
  float4 g = *reinterpret_cast<const float4*>(atw + off + q.i);
  // D_x^T p: p_x(i - 1) where x != 0, minus p_x(i) where row i has its -1
  {
    const float* px = p + off;
    const float4 c = *reinterpret_cast<const float4*>(px + q.i);
    const float prev = q.x > 0 ? px[q.i - 1] : 0.f;
    const bool last = q.x + 4 == v.nx;
    g.x += prev - c.x;
    g.y += c.x - c.y;
    g.z += c.y - c.z;
    g.w += c.z - ((v.mode != 0 || !last) ? c.w : 0.f);
  }
#pragma unroll
  for (int a = 1; a < 3; ++a) {
    const float* pa = p + a * plane + off;
    const float4 c = *reinterpret_cast<const float4*>(pa + q.i);
    if (quad_prev(v, q, a)) {
      const float4 b = *reinterpret_cast<const float4*>(pa + q.i - (unsigned)stride(v, a));
      g.x += b.x;
      g.y += b.y;
      g.z += b.z;
      g.w += b.w;
    }
    if (v.mode != 0 || quad_next(v, q, a)) {
      g.x -= c.x;
      g.y -= c.y;
      g.z -= c.z;
      g.w -= c.w;
    }
  }
  float4* fp = reinterpret_cast<float4*>(f + off + q.i);
  const float4 fo = *fp;
  const float4 fn = make_float4(fmaxf(fo.x - tau * g.x, 0.f), fmaxf(fo.y - tau * g.y, 0.f),
                                fmaxf(fo.z - tau * g.z, 0.f), fmaxf(fo.w - tau * g.w, 0.f));
  *fp = fn;
  *reinterpret_cast<float4*>(fbar + off + q.i) =
      make_float4(2.f * fn.x - fo.x, 2.f * fn.y - fo.y, 2.f * fn.z - fo.z, 2.f * fn.w - fo.w);
}

// checks shared by the single-volume entry points; v = the volume's shape
int vol_args(const int32_t* volume_size, int32_t batch, int32_t mode, Vol3& v) {
  if (!volume_size) return DTV_ERR_NULL_POINTER;
  if (batch < 1 || batch > 65535 || (mode != 0 && mode != 1)) return DTV_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 3; ++i)
    if (volume_size[i] < 1) return DTV_ERR_INVALID_ARGUMENT;
  if (volume_size[2] > 65535) return DTV_ERR_TOO_LARGE;
  v = make_vol3(volume_size[0], volume_size[1], volume_size[2], mode);
  if ((int64_t)v.tiles_x * ((v.ny + kTileY - 1) / kTileY) > 0x7fffffff) return DTV_ERR_TOO_LARGE;
  return DTV_OK;
}

}  // namespace

extern "C" {

int dtv_tv_fista_step(const float* y, const float* g, const float* xprev, float* xnew,
                      float* ynew, int32_t batch, const int32_t volume_size[3], float step,
                      float lam, float eps, float beta, int32_t nonneg, int32_t mode,
                      void* stream) {
  if (!y || !g || !xprev || !xnew || !ynew) return DTV_ERR_NULL_POINTER;
  // y is read by the neighbours' stencils: neither output may overwrite it
  if ((const float*)xnew == y || (const float*)ynew == y || xnew == ynew)
    return DTV_ERR_INVALID_ARGUMENT;
  Vol3 v;
  int rc = vol_args(volume_size, batch, mode, v);
  if (rc) return rc;
  dtv_tv_item it{y, g, xprev, xnew, ynew, volume_size[0], volume_size[1], volume_size[2],
                 nonneg ? 1 : 0, step, lam, eps, beta};
  const bool al = ((uintptr_t)y | (uintptr_t)g | (uintptr_t)xprev | (uintptr_t)xnew |
                   (uintptr_t)ynew) % 16 == 0;
  if (march4_ok(v) && al)
    hipLaunchKernelGGL(k_tv_fista_march4, dim3((unsigned)march4_blocks(v), batch), dim3(512), 0,
                       (hipStream_t)stream, it, v);
  else if (march_ok(v))
    hipLaunchKernelGGL(k_tv_fista_march, dim3((unsigned)march_blocks(v), batch), dim3(512), 0,
                       (hipStream_t)stream, it, v);
  else
    hipLaunchKernelGGL(k_tv_fista_one, vol3_grid(v, batch), dim3(256), 0, (hipStream_t)stream,
                       it, v);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

int dtv_tv_fista_step_multi(const dtv_tv_item* items_host, const dtv_tv_item* items_dev,
                            int32_t n_items, int32_t mode, void* stream) {
  if (!items_host || !items_dev) return DTV_ERR_NULL_POINTER;
  if (n_items < 1 || n_items > kMaxTvItems || (mode != 0 && mode != 1))
    return DTV_ERR_INVALID_ARGUMENT;
  TvOffsets offs;
  offs.n = n_items;
  bool march = true;
  for (int k = 0; k < n_items; ++k) {
    const dtv_tv_item& it = items_host[k];
    if (!it.y || !it.g || !it.xprev || !it.xnew || !it.ynew) return DTV_ERR_NULL_POINTER;
    if (it.nx < 1 || it.ny < 1 || it.nz < 1) return DTV_ERR_INVALID_ARGUMENT;
    if ((const float*)it.xnew == it.y || (const float*)it.ynew == it.y || it.xnew == it.ynew)
      return DTV_ERR_INVALID_ARGUMENT;
    march = march && march_ok(make_vol3(it.nx, it.ny, it.nz, mode));
  }
  int64_t total = 0;
  for (int k = 0; k < n_items; ++k) {
    const dtv_tv_item& it = items_host[k];
    const Vol3 v = make_vol3(it.nx, it.ny, it.nz, mode);
    offs.off[k] = (int)total;
    total += march ? (int64_t)march_blocks(v)
                   : (int64_t)v.tiles_x * ((v.ny + kTileY - 1) / kTileY) * v.nz;
    if (total > 0x7fffffff) return DTV_ERR_TOO_LARGE;
  }
  offs.off[n_items] = (int)total;
  if (march)
    hipLaunchKernelGGL(k_tv_fista_march_multi, dim3((unsigned)total), dim3(512), 0,
                       (hipStream_t)stream, items_dev, mode, offs);
  else
    hipLaunchKernelGGL(k_tv_fista_multi, dim3((unsigned)total), dim3(256), 0,
                       (hipStream_t)stream, items_dev, mode, offs);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

int dtv_tv_descent(const float* f, float* out, int32_t batch, const int32_t volume_size[3],
                   float step, float eps, int32_t mode, void* stream) {
  if (!f || !out) return DTV_ERR_NULL_POINTER;
  Vol3 v;
  int rc = vol_args(volume_size, batch, mode, v);
  if (rc) return rc;
  if (!(eps > 0.f) || f == out) return DTV_ERR_INVALID_ARGUMENT;
  if (march4_ok(v) && ((uintptr_t)f | (uintptr_t)out) % 16 == 0)
    hipLaunchKernelGGL(k_tv_descent_march4, dim3((unsigned)march4_blocks(v), batch), dim3(512),
                       0, (hipStream_t)stream, f, out, v, step, eps);
  else if (march_ok(v))
    hipLaunchKernelGGL(k_tv_descent_march, dim3((unsigned)march_blocks(v), batch), dim3(512), 0,
                       (hipStream_t)stream, f, out, v, step, eps);
  else
    hipLaunchKernelGGL(k_tv_descent, vol3_grid(v, batch), dim3(256), 0, (hipStream_t)stream, f,
                       out, v, step, eps);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

int dtv_pdhg_dtv_step(float* f, float* fbar, const float* atw, float* p, int32_t batch,
                      const int32_t volume_size[3], float sigma, float tau, const float lam[3],
                      int32_t mode, void* stream) {
  if (!f || !fbar || !atw || !p || !lam) return DTV_ERR_NULL_POINTER;
  Vol3 v;
  int rc = vol_args(volume_size, batch, mode, v);
  if (rc) return rc;
  if ((const void*)f == fbar || (const void*)f == atw || (const void*)fbar == atw ||
      (const void*)p == f || (const void*)p == fbar || (const void*)p == atw)
    return DTV_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 3; ++i)
    if (!(lam[i] >= 0.f)) return DTV_ERR_INVALID_ARGUMENT;
  const int64_t plane = v.n * batch;
  hipStream_t st = (hipStream_t)stream;
  // float4 access needs 16-byte aligned bases as well
  const bool aligned = ((uintptr_t)f | (uintptr_t)fbar | (uintptr_t)atw | (uintptr_t)p) % 16 == 0;
  if (quad_ok(v) && aligned) {
    const dim3 grid = quad_grid(v, batch);
    hipLaunchKernelGGL(k_pdhg_dual4, grid, dim3(256), 0, st, fbar, p, v, plane, sigma, lam[0],
                       lam[1], lam[2]);
    if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
    hipLaunchKernelGGL(k_pdhg_primal4, grid, dim3(256), 0, st, f, fbar, atw, p, v, plane, tau);
    return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
  }
  const dim3 grid = vol3_grid(v, batch);
  hipLaunchKernelGGL(k_pdhg_dual, grid, dim3(256), 0, st, fbar, p, v, plane, sigma, lam[0],
                     lam[1], lam[2]);
  if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
  hipLaunchKernelGGL(k_pdhg_primal, grid, dim3(256), 0, st, f, fbar, atw, p, v, plane, tau);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

}  // extern "C"
