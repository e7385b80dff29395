// Fused shrinkage steps of the reference unrolls (gfx950).
//
// * l1-ball projection (model/PDHG/Utils/L1ball.py:4-17), for up to
//   DTV_L1_MAX_ITEMS vectors in one call.  The reference's damped Newton
//   iteration on lam for  S(lam) = sum_i relu(a_i - lam) = b  runs on the
//   device with no host synchronisation: every Newton step is one launch
//   whose workgroups reduce S and C = #{a_i > lam} over a grid-stride slice
//   (FP32 terms a_i - lam as the reference forms them, FP64 sums) and whose
//   last workgroup to finish (an atomic counter between a release and an
//   acquire fence) adds the partial sums in a fixed order and takes the step
//     lam <- lam - (S - b) / (0.001 - C)
//   while |S - b| > tol and fewer than max_iter steps were taken (the same
//   test, order and step count as the reference's loop; its first step is
//   taken from the sums at lam = 0 that also decide sum a_i < b, "inside").
//   Launches after convergence return at once.  A last launch writes
//     mode 0 (projection):  w = inside ? v : sign(v) relu(|v| - lam)
//     mode 1 (dual update of model/PDHG/DTVCP.py:44-46, a = |v| / sigma):
//                           w = v - sigma sign(v) (inside ? a : relu(a - lam))
//   State and partial sums live in a caller-provided workspace.
// * soft threshold with FISTA momentum (model/FISTA/FISTA.py:21-25):
//     d = y - step g;  x = sign(d) relu(|d| - thr);  y' = x + beta (x - x_prev)
// * one step of the directional-TV dual-gradient unroll
//   (model/FISTA/DTVFISTA.py:29-41):
//     z = relu(f - sigma g);  m_a = max over every voxel of |D_a z| - v_a
//     (torch.max over the whole tensor: one scalar per axis);
//     p_a = sign(D_a z) m_a;  f_new = z + sum_a t_a D_a^T p_a;
//     f_out = f_new + beta (f_new - f)
//   as two launches (the maxima, then the update), z and D_a z formed on
//   the fly from f and g, so one step reads f and g twice and writes f_out.
#include <hip/hip_runtime.h>

#include <math.h>

#include <mutex>

#include "dtv_stencil.h"
#include "dtvproj.h"

using namespace dtv;

namespace {

constexpr int kL1Blocks = 512;  // workgroups per vector and Newton step (upper bound)

struct L1State {
  double lam;
  int active;  // another Newton step is due
  int inside;  // sum_i a_i < b: the vector itself is the answer
  int iters;
  int phase;   // 0 until the sums at lam = 0 are in
  unsigned count;
  int pad;
};

struct L1Args {
  dtv_l1_item it[DTV_L1_MAX_ITEMS];
  int n;
  int mode;
  int max_iter;
  double tol;
};

__device__ __forceinline__ float l1_a(const L1Args& A, const dtv_l1_item& it, float v) {
  return A.mode == 0 ? fabsf(v) : fabsf(v) / it.sigma;
}

__global__ void k_l1_init(L1State* st, int n) {
  const int k = threadIdx.x;
  if (k >= n) return;
  L1State s{};
  s.lam = 0.0;
  s.active = 1;
  s.phase = 0;
  st[k] = s;
}

// Block-wide FP64 sum of (s, c) over 256 lanes; the result in lane 0.
__device__ __forceinline__ void block_sum2(double& s, double& c, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off, 64);
    c += __shfl_xor(c, off, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
    sh[2 * w] = s;
    sh[2 * w + 1] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s = 0.0;
    c = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      s += sh[2 * q];
      c += sh[2 * q + 1];
    }
  }
}

// One Newton step (or the first evaluation) of every still-active vector;
// blockIdx.y = vector.
__global__ __launch_bounds__(256) void k_l1_step(L1Args A, L1State* __restrict__ st,
                                                 double* __restrict__ part) {
  const int k = blockIdx.y;
  L1State* s = st + k;
  if (!s->active) return;
  const dtv_l1_item it = A.it[k];
  const float lamf = (float)s->lam;
  double S = 0.0, C = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < it.n; i += stride) {
    const float r = l1_a(A, it, it.v[i]) - lamf;
    if (r > 0.f) {
      S += (double)r;
      C += 1.0;
    }
  }
  __shared__ double sh[8];
  __shared__ int last;
  block_sum2(S, C, sh);
  double* pk = part + (size_t)k * 2 * kL1Blocks;
  if (threadIdx.x == 0) {
    pk[2 * blockIdx.x] = S;
    pk[2 * blockIdx.x + 1] = C;
    __threadfence();
    const unsigned done = atomicAdd(&s->count, 1u);
    last = done == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  // the last workgroup: every partial sum is visible after the acquire; the
  // lanes add fixed strided subsets, then a fixed tree (deterministic)
  __threadfence();
  S = 0.0;
  C = 0.0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
    S += __hip_atomic_load(&pk[2 * b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    C += __hip_atomic_load(&pk[2 * b + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();  // sh is reused
  block_sum2(S, C, sh);
  if (threadIdx.x != 0) return;
  s->count = 0u;
  const double obj = S - it.radius;
  if (s->phase == 0) {
    s->phase = 1;
    if (S < it.radius) {  // L1ball.py:5-6
      s->inside = 1;
      s->active = 0;
      return;
    }
    if (!(fabs(obj) > A.tol) || A.max_iter <= 0) {
      s->active = 0;
      return;
    }
  }
  // L1ball.py:10-13: the step from the sums at the current lam
  s->lam = s->lam - obj / (0.001 - C);
  s->iters += 1;
  s->active = fabs(obj) > A.tol && s->iters < A.max_iter;
}

__global__ __launch_bounds__(256) void k_l1_apply(L1Args A, const L1State* __restrict__ st) {
  const int k = blockIdx.y;
  const dtv_l1_item it = A.it[k];
  const L1State s = st[k];
  const float lamf = (float)s.lam;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < it.n; i += stride) {
    const float v = it.v[i];
    const float sg = v > 0.f ? 1.f : (v < 0.f ? -1.f : 0.f);
    if (A.mode == 0) {
      it.out[i] = s.inside ? v : sg * fmaxf(fabsf(v) - lamf, 0.f);
    } else {
      const float a = fabsf(v) / it.sigma;
      const float l = s.inside ? a : fmaxf(a - lamf, 0.f);
      it.out[i] = v - it.sigma * sg * l;
    }
  }
}

// ---- soft threshold + momentum -----------------------------------------------

__global__ __launch_bounds__(256) void k_shrink_step(const float* __restrict__ y,
                                                     const float* __restrict__ g,
                                                     const float* xprev, float* xnew,
                                                     float* ynew, int64_t n, float step,
                                                     float thr, float beta) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float d = y[i] - step * g[i];
    const float sg = d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f);
    const float x = sg * fmaxf(fabsf(d) - thr, 0.f);
    const float xp = xprev[i];  // read before writing: x_prev may be xnew
    xnew[i] = x;
    ynew[i] = x + beta * (x - xp);
  }
}

// ---- directional-TV dual-gradient step ---------------------------------------

__device__ __forceinline__ float zval(const float* __restrict__ f, const float* __restrict__ g,
                                      int64_t i, float sigma) {
  return fmaxf(f[i] - sigma * g[i], 0.f);
}

// D_a z at p, z = relu(f - sigma g)
__device__ __forceinline__ float zdiff(const Vol3& v, const float* __restrict__ f,
                                       const float* __restrict__ g, const Pos& p, int a,
                                       float sigma, float zi) {
  if (has_next(v, p, a)) return zval(f, g, p.i + stride(v, a), sigma) - zi;
  return v.mode == 0 ? 0.f : -zi;
}

// order-preserving map of floats to unsigned ints (for atomicMax)
__device__ __forceinline__ unsigned ord(float x) {
  const unsigned u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(unsigned u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_fdpg_init(unsigned* mx) {
  if (threadIdx.x < 3) mx[threadIdx.x] = ord(-INFINITY);
}

__global__ __launch_bounds__(256) void k_fdpg_max(const float* __restrict__ f,
                                                  const float* __restrict__ g, Vol3 v,
                                                  float sigma, float v0, float v1, float v2,
                                                  unsigned* __restrict__ mx) {
  Pos p;
  const bool in = vol3_pos(v, p);
  const float vv[3] = {v0, v1, v2};
  float m[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (in) {
    const int64_t off = (int64_t)blockIdx.z * v.n;
    const float* fb = f + off;
    const float* gb = g + off;
    const float zi = zval(fb, gb, p.i, sigma);
#pragma unroll
    for (int a = 0; a < 3; ++a) m[a] = fabsf(zdiff(v, fb, gb, p, a, sigma, zi)) - vv[a];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m[a] = fmaxf(m[a], __shfl_xor(m[a], off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a)
      if (m[a] > -INFINITY) atomicMax(&mx[a], ord(m[a]));
  }
}

__global__ __launch_bounds__(256) void k_fdpg_apply(const float* __restrict__ f,
                                                    const float* __restrict__ g,
                                                    float* __restrict__ fout, Vol3 v,
                                                    float sigma, float t0, float t1, float t2,
                                                    float beta,
                                                    const unsigned* __restrict__ mx) {
  Pos p;
  if (!vol3_pos(v, p)) return;
  const int64_t off = (int64_t)blockIdx.z * v.n;
  const float* fb = f + off;
  const float* gb = g + off;
  const float tt[3] = {t0, t1, t2};
  const float zi = zval(fb, gb, p.i, sigma);
  float fn = zi;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float m = unord(mx[a]);
    float dt = 0.f;
    if (has_prev(v, p, a)) {
      // row i - s_a has its +1 at i: D_a z there is z(i) - z(i - s_a)
      const float dp = zi - zval(fb, gb, p.i - stride(v, a), sigma);
      dt += (dp > 0.f ? 1.f : (dp < 0.f ? -1.f : 0.f)) * m;
    }
    if (has_self(v, p, a)) {
      const float d = zdiff(v, fb, gb, p, a, sigma, zi);
      dt -= (d > 0.f ? 1.f : (d < 0.f ? -1.f : 0.f)) * m;
    }
    fn += tt[a] * dt;
  }
  fout[off + p.i] = fn + beta * (fn - fb[p.i]);
}

unsigned flat_blocks(int64_t n) {
  const int64_t b = (n + 1023) / 1024;
  return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

// dtv_l1ball after the workspace checks: validates the items and enqueues
// the init, Newton-step and apply launches on st (ws: state + partial sums).
int run_l1ball(const dtv_l1_item* items, int32_t n_items, int32_t mode, double tol,
               int32_t max_iter, void* ws, hipStream_t st) {
  if (!items) return DTV_ERR_NULL_POINTER;
  if (n_items < 1 || n_items > DTV_L1_MAX_ITEMS || (mode != 0 && mode != 1) ||
      max_iter < 0 || max_iter > 100000 || !(tol >= 0.0))
    return DTV_ERR_INVALID_ARGUMENT;
  L1Args A{};
  int64_t nmax = 0;
  for (int k = 0; k < n_items; ++k) {
    const dtv_l1_item& it = items[k];
    if (!it.v || !it.out) return DTV_ERR_NULL_POINTER;
    if (it.n < 1 || !(it.radius >= 0.0)) return DTV_ERR_INVALID_ARGUMENT;
    if (mode == 1 && !(it.sigma > 0.f)) return DTV_ERR_INVALID_ARGUMENT;
    A.it[k] = it;
    nmax = it.n > nmax ? it.n : nmax;
  }
  A.n = n_items;
  A.mode = mode;
  A.max_iter = max_iter;
  A.tol = tol;
  L1State* state = (L1State*)ws;
  double* part = (double*)((char*)ws + sizeof(L1State) * DTV_L1_MAX_ITEMS);
  const unsigned nb = (unsigned)std::min<int64_t>(kL1Blocks, (nmax + 1023) / 1024);
  hipLaunchKernelGGL(k_l1_init, dim3(1), dim3(64), 0, st, state, n_items);
  if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
  const int steps = max_iter > 0 ? max_iter : 1;
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(k_l1_step, dim3(nb, n_items), dim3(256), 0, st, A, state, part);
    if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
  }
  hipLaunchKernelGGL(k_l1_apply, dim3(std::min<unsigned>(flat_blocks(nmax), 4096), n_items),
                     dim3(256), 0, st, A, state);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

}  // namespace

// ---- allocation-free l1-ball entry point --------------------------------------
//
// dtv_l1ball_noalloc keeps the Newton state and the partial sums in this
// file-scope device array instead of a caller workspace: room for
// DTV_L1_MAX_ITEMS vectors of any length (the reduction uses at most
// kL1Blocks workgroups per vector), 131,584 bytes, one copy per device for
// the life of the process.  Host code must not pass the array's name to a
// kernel: its device address comes from hipGetSymbolAddress, looked up once
// per device with that device current.
static __device__ __align__(256) unsigned char scratch[sizeof(L1State) * DTV_L1_MAX_ITEMS +
                                                       sizeof(double) * 2 * kL1Blocks *
                                                           DTV_L1_MAX_ITEMS];
static_assert(sizeof(scratch) == 131584, "the documented size of the l1-ball scratch array");

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices];

// The current device's address of `scratch`.
int scratch_address(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return DTV_ERR_HIP;
  if (dev < 0 || dev >= kMaxDevices) return DTV_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return DTV_ERR_HIP;
    }
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return DTV_OK;
}

}  // namespace

extern "C" {

int dtv_l1ball_noalloc(const dtv_l1_item* items, int32_t n_items, int32_t mode, double tol,
                       int32_t max_iter, void* stream) {
  if (n_items > DTV_L1_MAX_ITEMS) return DTV_ERR_TOO_LARGE;  // beyond the scratch array
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc) return rc;
  return run_l1ball(items, n_items, mode, tol, max_iter, ws, (hipStream_t)stream);
}

size_t dtv_l1ball_workspace_bytes(int32_t n_items) {
  if (n_items < 1 || n_items > DTV_L1_MAX_ITEMS) return 0;
  return sizeof(L1State) * DTV_L1_MAX_ITEMS +
         sizeof(double) * 2 * (size_t)kL1Blocks * (size_t)n_items;
}

int dtv_l1ball(const dtv_l1_item* items, int32_t n_items, int32_t mode, double tol,
               int32_t max_iter, void* workspace, size_t workspace_bytes, void* stream) {
  if (!workspace) return DTV_ERR_NULL_POINTER;
  if (n_items >= 1 && n_items <= DTV_L1_MAX_ITEMS &&
      workspace_bytes < dtv_l1ball_workspace_bytes(n_items))
    return DTV_ERR_INVALID_ARGUMENT;
  if ((uintptr_t)workspace % 8 != 0) return DTV_ERR_INVALID_ARGUMENT;
  return run_l1ball(items, n_items, mode, tol, max_iter, workspace, (hipStream_t)stream);
}

int dtv_shrink_step(const float* y, const float* g, const float* xprev, float* xnew,
                    float* ynew, int64_t n, float step, float thr, float beta, void* stream) {
  if (!y || !g || !xprev || !xnew || !ynew) return DTV_ERR_NULL_POINTER;
  if (n < 1 || (const void*)ynew == xprev || ynew == xnew) return DTV_ERR_INVALID_ARGUMENT;
  hipLaunchKernelGGL(k_shrink_step, dim3(flat_blocks(n)), dim3(256), 0, (hipStream_t)stream,
                     y, g, xprev, xnew, ynew, n, step, thr, beta);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

size_t dtv_fdpg_workspace_bytes(void) { return 64; }

int dtv_fdpg_step(const float* f, const float* g, float* fout, int32_t batch,
                  const int32_t volume_size[3], float sigma, const float v[3], const float t[3],
                  float beta, int32_t mode, void* workspace, size_t workspace_bytes,
                  void* stream) {
  if (!f || !g || !fout || !volume_size || !v || !t || !workspace) return DTV_ERR_NULL_POINTER;
  if (batch < 1 || batch > 65535 || (mode != 0 && mode != 1)) return DTV_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 3; ++i)
    if (volume_size[i] < 1) return DTV_ERR_INVALID_ARGUMENT;
  if (volume_size[2] > 65535) return DTV_ERR_TOO_LARGE;
  if ((const void*)fout == f || (const void*)fout == g) return DTV_ERR_INVALID_ARGUMENT;
  if (workspace_bytes < dtv_fdpg_workspace_bytes() || (uintptr_t)workspace % 4 != 0)
    return DTV_ERR_INVALID_ARGUMENT;
  const Vol3 vol = make_vol3(volume_size[0], volume_size[1], volume_size[2], mode);
  unsigned* mx = (unsigned*)workspace;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid = vol3_grid(vol, batch);
  hipLaunchKernelGGL(k_fdpg_init, dim3(1), dim3(64), 0, st, mx);
  if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
  hipLaunchKernelGGL(k_fdpg_max, grid, dim3(256), 0, st, f, g, vol, sigma, v[0], v[1], v[2], mx);
  if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
  hipLaunchKernelGGL(k_fdpg_apply, grid, dim3(256), 0, st, f, g, fout, vol, sigma, t[0], t[1],
                     t[2], beta, mx);
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

}  // extern "C"
