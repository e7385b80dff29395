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
// Work items are (instance, voxel) over a descriptor table, so uneven
// instances share one launch.
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

#include "dtvproj.h"

namespace {

struct Stencil {
  const float* __restrict__ f;
  int nx, ny, nz;
  int64_t n;
  int64_t s[3];
  int mode;
};

// d_a at flat index i (i must be a valid voxel)
__device__ __forceinline__ float diff(const Stencil& S, int64_t i, int a) {
  const float fi = S.f[i];
  if (S.mode == 0) {
    bool ok;
    if (a == 0) ok = (i % S.nx) != S.nx - 1;
    else if (a == 1) ok = ((i / S.nx) % S.ny) != S.ny - 1;
    else ok = (i / ((int64_t)S.nx * S.ny)) != S.nz - 1;
    return ok ? S.f[i + S.s[a]] - fi : 0.f;
  }
  bool ok;
  if (a == 0) ok = ((i + 1) % S.nx) != 0;
  else ok = i < S.n - S.s[a];
  return ok ? S.f[i + S.s[a]] - fi : -fi;
}

// does row i - s_a of D_a carry a +1 at column i?
__device__ __forceinline__ bool has_prev(const Stencil& S, int64_t i, int a) {
  if (S.mode == 0) {
    if (a == 0) return (i % S.nx) != 0;
    if (a == 1) return ((i / S.nx) % S.ny) != 0;
    return i / ((int64_t)S.nx * S.ny) != 0;
  }
  if (a == 0) return (i % S.nx) != 0;
  return i >= S.s[a];
}

// q = D f / n at voxel i: n = sqrt(|D f|^2 + eps^2) (kAsd false) or
// |D f| + eps (kAsd true)
template <bool kAsd>
__device__ __forceinline__ void qvec(const Stencil& S, int64_t i, float eps, float q[3]) {
  const float dx = diff(S, i, 0), dy = diff(S, i, 1), dz = diff(S, i, 2);
  const float ss = dx * dx + dy * dy + dz * dz;
  const float r = kAsd ? 1.f / (sqrtf(ss) + eps) : rsqrtf(ss + eps * eps);
  q[0] = dx * r;
  q[1] = dy * r;
  q[2] = dz * r;
}

__device__ __forceinline__ Stencil make_stencil(const float* f, int nx, int ny, int nz,
                                                int mode) {
  Stencil S;
  S.f = f;
  S.nx = nx;
  S.ny = ny;
  S.nz = nz;
  S.n = (int64_t)nx * ny * nz;
  S.s[0] = 1;
  S.s[1] = mode == 0 ? nx : ny;
  S.s[2] = (int64_t)nx * ny;
  S.mode = mode;
  return S;
}

// sum_a D_a^T q_a at voxel i
template <bool kAsd>
__device__ __forceinline__ float tv_grad_at(const Stencil& S, int64_t i, float eps) {
  float q[3];
  qvec<kAsd>(S, i, eps, q);
  float gtv = -(q[0] + q[1] + q[2]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (has_prev(S, i, a)) {
      float qp[3];
      qvec<kAsd>(S, i - S.s[a], eps, qp);
      gtv += qp[a];
    }
  }
  return gtv;
}

__device__ __forceinline__ void tv_point(const dtv_tv_item& it, int mode, int64_t off,
                                         int64_t i) {
  const Stencil S = make_stencil(it.y + off, it.nx, it.ny, it.nz, mode);
  const float gtv = tv_grad_at<false>(S, i, it.eps);
  // read every input at i before writing: x_prev may share storage with an
  // output (y may not: neighbours read it)
  const float yv = S.f[i];
  const float gv = it.g[off + i];
  const float xp = it.xprev[off + i];
  float x = yv - it.step * (gv + it.lam * gtv);
  if (it.nonneg) x = fmaxf(x, 0.f);
  it.xnew[off + i] = x;
  it.ynew[off + i] = x + it.beta * (x - xp);
}

// one instance, blockIdx.y = batch item (volumes of equal size, contiguous)
__global__ __launch_bounds__(256) void k_tv_fista_one(dtv_tv_item it, int mode) {
  const int64_t nvox = (int64_t)it.nx * it.ny * it.nz;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvox) return;
  tv_point(it, mode, (int64_t)blockIdx.y * nvox, i);
}

// uneven instances from a device descriptor table, blockIdx.y = instance
__global__ __launch_bounds__(256) void k_tv_fista_multi(const dtv_tv_item* __restrict__ items,
                                                        int mode) {
  const dtv_tv_item it = items[blockIdx.y];
  const int64_t nvox = (int64_t)it.nx * it.ny * it.nz;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvox) return;
  tv_point(it, mode, 0, i);
}

// out = f - step * TV gradient (ASD normaliser), blockIdx.y = batch item
__global__ __launch_bounds__(256) void k_tv_descent(const float* __restrict__ f,
                                                    float* __restrict__ out, int nx, int ny,
                                                    int nz, float step, float eps, int mode) {
  const int64_t nvox = (int64_t)nx * ny * nz;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvox) return;
  const int64_t off = (int64_t)blockIdx.y * nvox;
  const Stencil S = make_stencil(f + off, nx, ny, nz, mode);
  out[off + i] = S.f[i] - step * tv_grad_at<true>(S, i, eps);
}

// one PDHG step of the directional-TV unroll, after A and A^T.  Dual half:
//   p_a <- clamp(p_a + sigma * d_a(fbar), -lam_a, lam_a)
// (p: three planes [3][batch][Z][Y][X], updated in place; each lane reads
// only its own p entries).
__global__ __launch_bounds__(256) void k_pdhg_dual(const float* __restrict__ fbar,
                                                   float* __restrict__ p, int nx, int ny,
                                                   int nz, int64_t plane, float sigma, float l0,
                                                   float l1, float l2, int mode) {
  const int64_t nvox = (int64_t)nx * ny * nz;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvox) return;
  const int64_t off = (int64_t)blockIdx.y * nvox;
  const Stencil S = make_stencil(fbar + off, nx, ny, nz, mode);
  const float lam[3] = {l0, l1, l2};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float* pa = p + a * plane + off;
    const float v = pa[i] + sigma * diff(S, i, a);
    pa[i] = fminf(fmaxf(v, -lam[a]), lam[a]);
  }
}

// Primal half: g = atw + sum_a D_a^T p_a;  f_new = max(f - tau g, 0);
// fbar <- 2 f_new - f;  f <- f_new.  D_a^T p at i is p_a(i - s_a) where row
// i - s_a of D_a has its +1 at i, minus p_a(i) where row i is not empty (in
// the reference boundary mode every row has its -1).
__global__ __launch_bounds__(256) void k_pdhg_primal(float* __restrict__ f,
                                                     float* __restrict__ fbar,
                                                     const float* __restrict__ atw,
                                                     const float* __restrict__ p, int nx,
                                                     int ny, int nz, int64_t plane, float tau,
                                                     int mode) {
  const int64_t nvox = (int64_t)nx * ny * nz;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nvox) return;
  const int64_t off = (int64_t)blockIdx.y * nvox;
  const Stencil S = make_stencil(nullptr, nx, ny, nz, mode);
  float g = atw[off + i];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float* pa = p + a * plane + off;
    float dt = has_prev(S, i, a) ? pa[i - S.s[a]] : 0.f;
    bool own;
    if (mode != 0) own = true;
    else if (a == 0) own = (i % nx) != nx - 1;
    else if (a == 1) own = ((i / nx) % ny) != ny - 1;
    else own = (i / ((int64_t)nx * ny)) != nz - 1;
    if (own) dt -= pa[i];
    g += dt;
  }
  const float fo = f[off + i];
  const float fn = fmaxf(fo - tau * g, 0.f);
  f[off + i] = fn;
  fbar[off + i] = 2.f * fn - fo;
}

}  // namespace

extern "C" {

int dtv_tv_fista_step(const float* y, const float* g, const float* xprev, float* xnew,
                      float* ynew, int32_t batch, const int32_t volume_size[3], float step,
                      float lam, float eps, float beta, int32_t nonneg, int32_t mode,
                      void* stream) {
  if (!y || !g || !xprev || !xnew || !ynew || !volume_size) return DTV_ERR_NULL_POINTER;
  if (batch < 1 || batch > 65535 || (mode != 0 && mode != 1)) return DTV_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 3; ++i)
    if (volume_size[i] < 1) return DTV_ERR_INVALID_ARGUMENT;
  dtv_tv_item it{y, g, xprev, xnew, ynew, volume_size[0], volume_size[1], volume_size[2],
                 nonneg ? 1 : 0, step, lam, eps, beta};
  const int64_t nvox = (int64_t)volume_size[0] * volume_size[1] * volume_size[2];
  const int64_t blocks = (nvox + 255) / 256;
  if (blocks > 0x7fffffff) return DTV_ERR_TOO_LARGE;
  hipLaunchKernelGGL(k_tv_fista_one, dim3((unsigned)blocks, batch), dim3(256), 0,
                     (hipStream_t)stream, it, mode);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

int dtv_tv_fista_step_multi(const dtv_tv_item* items, int32_t n_items, int64_t max_voxels,
                            int32_t mode, void* stream) {
  if (!items) return DTV_ERR_NULL_POINTER;
  if (n_items < 1 || n_items > 65535 || max_voxels < 1 || (mode != 0 && mode != 1))
    return DTV_ERR_INVALID_ARGUMENT;
  const int64_t blocks = (max_voxels + 255) / 256;
  if (blocks > 0x7fffffff) return DTV_ERR_TOO_LARGE;
  hipLaunchKernelGGL(k_tv_fista_multi, dim3((unsigned)blocks, n_items), dim3(256), 0,
                     (hipStream_t)stream, items, mode);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

int dtv_tv_descent(const float* f, float* out, int32_t batch, const int32_t volume_size[3],
                   float step, float eps, int32_t mode, void* stream) {
  if (!f || !out || !volume_size) return DTV_ERR_NULL_POINTER;
  if (batch < 1 || batch > 65535 || (mode != 0 && mode != 1) || !(eps > 0.f) || f == out)
    return DTV_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 3; ++i)
    if (volume_size[i] < 1) return DTV_ERR_INVALID_ARGUMENT;
  const int64_t nvox = (int64_t)volume_size[0] * volume_size[1] * volume_size[2];
  const int64_t blocks = (nvox + 255) / 256;
  if (blocks > 0x7fffffff) return DTV_ERR_TOO_LARGE;
  hipLaunchKernelGGL(k_tv_descent, dim3((unsigned)blocks, batch), dim3(256), 0,
                     (hipStream_t)stream, f, out, volume_size[0], volume_size[1],
                     volume_size[2], step, eps, mode);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

int dtv_pdhg_dtv_step(float* f, float* fbar, const float* atw, float* p, int32_t batch,
                      const int32_t volume_size[3], float sigma, float tau, const float lam[3],
                      int32_t mode, void* stream) {
  if (!f || !fbar || !atw || !p || !volume_size || !lam) return DTV_ERR_NULL_POINTER;
  if (batch < 1 || batch > 65535 || (mode != 0 && mode != 1))
    return DTV_ERR_INVALID_ARGUMENT;
  if ((const void*)f == fbar || (const void*)f == atw || (const void*)fbar == atw ||
      (const void*)p == f || (const void*)p == fbar || (const void*)p == atw)
    return DTV_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 3; ++i)
    if (volume_size[i] < 1 || !(lam[i] >= 0.f)) return DTV_ERR_INVALID_ARGUMENT;
  const int64_t nvox = (int64_t)volume_size[0] * volume_size[1] * volume_size[2];
  const int64_t blocks = (nvox + 255) / 256;
  if (blocks > 0x7fffffff) return DTV_ERR_TOO_LARGE;
  const int64_t plane = nvox * batch;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pdhg_dual, dim3((unsigned)blocks, batch), dim3(256), 0, st, fbar, p,
                     volume_size[0], volume_size[1], volume_size[2], plane, sigma, lam[0],
                     lam[1], lam[2], mode);
  if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
  hipLaunchKernelGGL(k_pdhg_primal, dim3((unsigned)blocks, batch), dim3(256), 0, st, f, fbar,
                     atw, p, volume_size[0], volume_size[1], volume_size[2], plane, tau, mode);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

}  // extern "C"
