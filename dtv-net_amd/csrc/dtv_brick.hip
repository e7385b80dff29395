// Volume-stationary ("brick") kernels for A and the matched A^T on gfx950.
//
// One workgroup owns one brick of the volume for the whole call and visits
// every view.  Per view it projects the 8 corners of the brick's sample
// support box onto the detector (continuous pixel coordinates), takes the
// bounding rectangle, and lets each lane march one ray of that rectangle
// through the brick using the same fixed-point positions as the ray-driven
// kernels (shared ray_core, so A and A^T stay exact transposes).
//
//   k_fwd_brick  A: the brick plus a one-voxel halo (zero outside the
//                volume) sits in LDS; every sample belongs to exactly one
//                brick (the one holding floor(p)), so the 8 taps are plain
//                LDS reads.  The per-ray partial sum is added to the
//                sinogram with one hardware FP32 atomic per (ray, brick).
//                The volume is read from HBM once per call.
//   k_adj_brick  matched A^T: per view, each sample whose stencil touches
//                the brick adds its taps into fixed-point int32 LDS
//                accumulators (integer LDS atomics are several times faster
//                than FP32 ones on gfx950 and order-independent); after each
//                group of views sharing one scale the owned integers are
//                scaled back into per-lane FP32 registers.  The brick is
//                written to HBM once at the end: no global atomics, no
//                zero-fill, bitwise reproducible.
#include <hip/hip_runtime.h>

#include "dtv_geom.h"
#include "dtv_launch.h"
#include "dtvproj.h"

using namespace dtv;

// Brick shapes compiled with constant dimensions (the autotuner's candidates);
// any other shape runs the run-time-shape kernels.
#define DTV_BRICK_SHAPES(X) \
  X(32, 16, 16) X(26, 26, 14) X(24, 24, 14) X(32, 32, 8) X(16, 16, 32) X(16, 16, 16) X(32, 16, 8) \
  X(16, 16, 8) X(8, 8, 8)


namespace {

// ---- per-chunk view staging --------------------------------------------------
//
// Views are taken in chunks of kChunk.  A prologue stages a whole chunk at
// once, one (view, corner) pair per lane: each lane projects one corner of the
// brick's sample box onto its view's detector, the 8 lanes of a view reduce
// their corners with shuffles, and the first of them stores the view's
// rectangle of candidate pixels (the bounding box of the projected corners,
// rounded outwards; the whole detector when a corner is not in front of the
// source), its FP32 vectors for the pass-1 line test and, for A^T, the
// bound on one voxel's weight sum used to pick the fixed-point scale.
//
// Before the chunks, one lane per view tests whether the box can reach the
// view's detector at all (the same corner projections, with three pixels more
// slack than the staging's outward rounding); chunks are then formed from the
// views that can, in increasing view order.  On scans with many narrow sub-detectors most
// views never see a given brick, and they no longer cost a chunk slot.

constexpr int kChunk = 32;
constexpr int kCullWords = 128;  // views culled per brick: up to 32 * kCullWords

// Culling pays only when it can remove chunks: scans of more than four chunks
// (and at most 32 * kCullWords views) are culled, shorter ones are not.
__device__ __forceinline__ bool culls_views(const Geo& g) {
  return g.n_views > 4 * kChunk && g.n_views <= 32 * kCullWords;
}

struct ChunkStage {
  float fview[kChunk][12];  // S, D0, U, V in FP32
  int rect[kChunk][4];      // u0, u1, v0, v1 (inclusive; u1 < u0 = empty)
  float wbound[kChunk];     // A^T: bound on sum_k w for one voxel and view
  unsigned vmax[kChunk];    // A^T: largest |value| * scale over the rectangle (float bits)
  View sv[2];               // FP64 view of the current view (double-buffered)
  int vidx[kChunk];         // view index of each chunk slot
  unsigned live[kCullWords];  // views whose projected box can reach the detector
  int livepre[kCullWords];    // live views in the words before each word
  int n_live;                 // chunked views (all of them without culling)
};

// Marks in st.live the views that the box [lo, hi) can reach (see above) and
// sets st.livepre / st.n_live (only called when culls_views(g)).  A view is
// dropped only when all 8 corners project in front of the
// source and their bounding box, widened by 3 pixels, misses the detector, so
// its staged rectangle would have been empty.  Ends with a
// barrier.
__device__ void cull_views(const Geo& g, const double lo[3], const double hi[3],
                           ChunkStage& st) {
  for (int w = threadIdx.x; w < kCullWords; w += blockDim.x) st.live[w] = 0u;
  __syncthreads();
  {
    for (int a = threadIdx.x; a < g.n_views; a += blockDim.x) {
      View v;
      load_view(g, a, v);
      ViewProj vp;
      view_proj(v, vp);
      bool keep = false;
      double amin = 0.0, amax = 0.0, bmin = 0.0, bmax = 0.0;
      for (int t = 0; t < 8; ++t) {
        const double c[3] = {(t & 1) ? hi[0] : lo[0], (t & 2) ? hi[1] : lo[1],
                             (t & 4) ? hi[2] : lo[2]};
        double pa = 0.0, pb = 0.0;
        if (!project_point(v, vp, c, pa, pb)) keep = true;
        amin = t ? fmin(amin, pa) : pa;
        amax = t ? fmax(amax, pa) : pa;
        bmin = t ? fmin(bmin, pb) : pb;
        bmax = t ? fmax(bmax, pb) : pb;
      }
      // comparisons with NaN are false: such a view is kept
      const bool miss = ceil(amax) + 3.0 < 0.0 || floor(amin) - 3.0 > (double)(g.nu - 1) ||
                        ceil(bmax) + 3.0 < 0.0 || floor(bmin) - 3.0 > (double)(g.nv - 1);
      if (keep || !miss) atomicOr(&st.live[a >> 5], 1u << (a & 31));
    }
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    // exclusive prefix of the words' popcounts, two words per lane
    const int l = threadIdx.x;
    const int c0 = __popc(st.live[2 * l]), c1 = __popc(st.live[2 * l + 1]);
    int incl = c0 + c1;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(incl, off, 64);
      if (l >= off) incl += y;
    }
    st.livepre[2 * l] = incl - c0 - c1;
    st.livepre[2 * l + 1] = incl - c1;
    if (l == 63) st.n_live = incl;
  }
  __syncthreads();
}

// Index of the j-th chunked view: the j-th set bit of st.live, or j itself
// without culling.
__device__ __forceinline__ int chunk_view(const Geo& g, const ChunkStage& st, int j) {
  if (!culls_views(g)) return j;
  int w = 0;  // last word whose prefix is <= j
#pragma unroll
  for (int step = kCullWords / 2; step > 0; step >>= 1)
    if (st.livepre[w + step] <= j) w += step;
  unsigned m = st.live[w];
  int r = j - st.livepre[w], pos = 0;
#pragma unroll
  for (int sh = 16; sh > 0; sh >>= 1) {
    const unsigned low = m & ((1u << sh) - 1u);
    const int c = __popc(low);
    if (r >= c) {
      r -= c;
      m >>= sh;
      pos += sh;
    } else {
      m = low;
    }
  }
  return 32 * w + pos;
}

// A brick's views may be split over nparts workgroups (grid z): part p takes
// the chunked views [end(p - 1), end(p)), end(p) = n (p + 1) / nparts.
__device__ __forceinline__ int view_part_end(int n, int part, int nparts) {
  return (int)((long long)n * (part + 1) / nparts);
}

// Stages chunked views [c0, c0+n) of the box [lo, hi) (voxel frame) into st
// (their view indices into st.vidx).
__device__ void stage_chunk(const Geo& g, int c0, int n, const double lo[3], const double hi[3],
                            bool bound, ChunkStage& st) {
  for (int pi = threadIdx.x; pi < n * 8; pi += blockDim.x) {
    const int vi = pi >> 3, t = pi & 7;
    const int va = chunk_view(g, st, c0 + vi);
    View v;
    load_view(g, va, v);
    ViewProj vp;
    view_proj(v, vp);
    const double c[3] = {(t & 1) ? hi[0] : lo[0], (t & 2) ? hi[1] : lo[1],
                         (t & 4) ? hi[2] : lo[2]};
    double a = 0.0, b = 0.0;
    int ok = project_point(v, vp, c, a, b) ? 1 : 0;
    const double w[3] = {c[0] - v.S[0], c[1] - v.S[1], c[2] - v.S[2]};
    const double dw = fabs(dot3(w, vp.n));
    double amin = a, amax = a, bmin = b, bmax = b, wmin = dw, wmax = dw;
    // the 8 corners of a view are 8 consecutive lanes of one wave (n * 8
    // lanes take part, so every aligned group of 8 is whole); corner
    // t ^ (1 << k) is the other end of corner t's edge along axis k
    double du[3], dv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      du[k] = fabs(a - __shfl_xor(a, 1 << k, 64));
      dv[k] = fabs(b - __shfl_xor(b, 1 << k, 64));
    }
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        du[k] = fmax(du[k], __shfl_xor(du[k], off, 64));
        dv[k] = fmax(dv[k], __shfl_xor(dv[k], off, 64));
      }
      amin = fmin(amin, __shfl_xor(amin, off, 64));
      amax = fmax(amax, __shfl_xor(amax, off, 64));
      bmin = fmin(bmin, __shfl_xor(bmin, off, 64));
      bmax = fmax(bmax, __shfl_xor(bmax, off, 64));
      wmin = fmin(wmin, __shfl_xor(wmin, off, 64));
      wmax = fmax(wmax, __shfl_xor(wmax, off, 64));
      ok &= __shfl_xor(ok, off, 64);
    }
    if (t == 0) {
      int rect[4];
      float wb;
      box_view_finish(amin, amax, bmin, bmax, wmin, wmax, du, dv, ok != 0, lo, hi, g.nu, g.nv,
                      rect, wb);
      for (int k = 0; k < 4; ++k) st.rect[vi][k] = rect[k];
      st.vidx[vi] = va;
      if (bound) st.wbound[vi] = wb;
      st.vmax[vi] = 0u;
      for (int k = 0; k < 3; ++k) {
        st.fview[vi][k] = (float)v.S[k];
        st.fview[vi][3 + k] = (float)v.D0[k];
        st.fview[vi][6 + k] = (float)v.U[k];
        st.fview[vi][9 + k] = (float)v.V[k];
      }
    }
  }
}

// ---- fixed-point ray sums ------------------------------------------------------
//
// The deterministic forward adds each (ray, brick) partial sum into an int64
// per ray at the power-of-two scale S with S x bound < 2^62, bound = max|vol|
// x |output scale| x (samples a ray can have inside the volume's support box:
// one per unit of length along the ray, so its diagonal plus two).  Every
// partial and every running sum of a ray stays below that bound in magnitude
// (trilinear weights of a sample sum to at most 1), so nothing overflows, and
// integer addition makes the result independent of the order of the adds.
// Returns 0 for a non-finite max (the FP32 path is taken instead); an all-zero
// volume gets S = 1.
__device__ __forceinline__ double fix_ray_scale(unsigned absmax_bits, const Geo& g) {
  if (absmax_bits >= 0x7f800000u) return 0.0;
  const double dx = g.nx + 2.0, dy = g.ny + 2.0, dz = g.nz + 2.0;
  const double len = sqrt(dx * dx + dy * dy + dz * dz) + 2.0;
  const double bound = (double)__uint_as_float(absmax_bits) * fabs(g.scale) * len;
  if (!(bound > 0.0)) return 1.0;
  int e;
  frexp(bound, &e);  // bound < 2^e
  return ldexp(1.0, 62 - e);
}

// ---- ray scheduling ----------------------------------------------------------
//
// Per (brick, view) the rectangle of candidate rays is taken in segments of
// kSchedR rays per lane.  Pass 1 tests each candidate's line against the
// brick's sample box in FP32 with kLineSlack of slack (no ray that has a
// sample in the box can fail it) and estimates its chord in samples; the
// hits are counting-sorted by chord, longest first, into an LDS list.  Pass
// 2 marches the list: each wave gets rays of similar length, and rays that
// miss the brick cost no wave time.  Pass 2 solves every ray's run of
// samples in the brick exactly (exact_window), so its sample loop carries no
// membership test.

constexpr int kSchedR = 8;       // candidate rays per lane per segment (256-lane workgroups)
constexpr int kSchedBins = 64;   // chord classes, one per sample (the last open-ended)
constexpr int kSchedCap = kSchedR * 256;  // also 4 per lane of a 512-lane workgroup

struct RaySched {
  unsigned hist[kSchedBins];     // hits per class; cleared by the scan
  unsigned offs[kSchedBins];     // first list slot of each class
  unsigned n;                    // hits in this segment
  unsigned short list[kSchedCap];  // segment-relative candidate indices
};

// q = i / n, r = i % n for 0 <= i < 2^24 (float reciprocal rn = 1/n, one
// correction step: the estimate is within one of the quotient)
__device__ __forceinline__ void divmod(int i, int n, float rn, int& q, int& r) {
  q = (int)((float)i * rn);
  r = i - (int)__umul24((unsigned)q, (unsigned)n);  // q, n < 2^24
  if (r < 0) {
    --q;
    r += n;
  } else if (r >= n) {
    ++q;
    r -= n;
  }
}

// Wave-uniform copies of values read from LDS: readfirstlane puts them in
// scalar registers, so the per-ray set-up takes them as SGPR operands and the
// vector registers stay free for the sample loop and the A^T accumulators.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uni(double x) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ void uni_view(const View& src, View& v) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v.S[c] = uni(src.S[c]);
    v.D0[c] = uni(src.D0[c]);
    v.U[c] = uni(src.U[c]);
    v.V[c] = uni(src.V[c]);
  }
}

// Slack of the pass-1 box around a brick's sample box, in voxels.  The FP32
// line test is off by well under 1e-3 voxel for any geometry whose source
// lies within 1e4 voxels (rounded FP32 view vectors, one-ulp reciprocal); rays
// that only graze the box are dropped by the exact window of pass 2 anyway.
constexpr float kLineSlack = 0.0625f;

// Length in samples (one per voxel of path) of the pixel (iu, iv)'s line
// inside the box [blo, bhi], or -1 when the line misses it; fv = the view's
// FP32 vectors.  The callers' boxes carry kLineSlack, far above the rounding
// here.
__device__ __forceinline__ float chord_samples(const float* fv, int iu, int iv,
                                               const float blo[3], const float bhi[3]) {
  float tin = -3.0e38f, tout = 3.0e38f, rr = 0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float rc = fv[3 + c] + (float)iu * fv[6 + c] + (float)iv * fv[9 + c] - fv[c];
    rr += rc * rc;
    rc = rc != 0.f ? rc : 1e-30f;  // parallel: huge parameters, right sign of the test
    const float ir = __builtin_amdgcn_rcpf(rc);
    const float t0 = (blo[c] - fv[c]) * ir, t1 = (bhi[c] - fv[c]) * ir;
    tin = fmaxf(tin, fminf(t0, t1));
    tout = fminf(tout, fmaxf(t0, t1));
  }
  return tout >= tin ? (tout - tin) * sqrtf(rr) : -1.f;
}

// Pass-1 bookkeeping of one hit: class and rank within the class.
__device__ __forceinline__ unsigned sched_hit(RaySched& rs, float chord) {
  const unsigned key = (unsigned)min((int)chord, kSchedBins - 1);
  return (key << 16) | atomicAdd(&rs.hist[key], 1u);
}

// Counting sort of the segment's hits (slot[j] = class << 16 | rank, or ~0)
// into rs.list, longest class first; rs.n = number of hits.
template <int kR>
__device__ __forceinline__ void sched_sort(RaySched& rs, const unsigned (&slot)[kR]) {
  __syncthreads();
  if (threadIdx.x < kSchedBins) {
    const int l = threadIdx.x;
    const int key = kSchedBins - 1 - l;
    const unsigned h = rs.hist[key];
    unsigned incl = h;
#pragma unroll
    for (int off = 1; off < kSchedBins; off <<= 1) {
      const unsigned o = __shfl_up(incl, off, kSchedBins);
      if (l >= off) incl += o;
    }
    rs.offs[key] = incl - h;
    rs.hist[key] = 0u;
    if (l == kSchedBins - 1) rs.n = incl;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kR; ++j)
    if (slot[j] != ~0u)
      rs.list[rs.offs[slot[j] >> 16] + (slot[j] & 0xffffu)] =
          (unsigned short)(threadIdx.x + j * blockDim.x);
  __syncthreads();
}

// Pass 2 hands the sorted list (longest class first) to the waves in
// 64-ray blocks, one block per wave per round of blockDim.x rays, in snake
// order: wave w takes block w of even rounds and block nw - 1 - w of odd
// ones, so the wave with the longest rays of one round gets the shortest of
// the next and the waves reach the next barrier together (on the 256^3 /
// 60-view pair: A 2.76 -> 2.61 ms, A^T 2.97 -> 2.87 ms against the plain
// round-robin order).  Every list index below the hit count is visited once.
__device__ __forceinline__ int pass2_index(int p0, int odd) {
  const int w = (int)(threadIdx.x >> 6), nw = (int)(blockDim.x >> 6);
  return p0 + ((odd ? nw - 1 - w : w) << 6) + (int)(threadIdx.x & 63);
}

// Candidate index -> pixel.  Forward: row-major over the rectangle.
// Adjoint: the rectangle split into its nine stride-3 sub-lattices, so rays
// that end up side by side in a wave are mostly three pixels apart and
// rarely add into one voxel in the same instruction.
struct RectMap {
  int u0, v0, nur, nvr, mu, mv, w;  // w = row length of the index space
  float rw;
  int n;                            // index space size
  __device__ void init(const int* rect, bool interleave) {
    u0 = uni(rect[0]);
    v0 = uni(rect[2]);
    nur = uni(rect[1]) - u0 + 1;
    nvr = uni(rect[3]) - v0 + 1;
    mu = (nur + 2) / 3;
    mv = (nvr + 2) / 3;
    w = interleave ? 3 * mu : nur;
    n = interleave ? 9 * mu * mv : nur * nvr;
    rw = 1.f / (float)w;
  }
  // false for the padding of the interleaved space
  __device__ __forceinline__ bool pixel(int i, bool interleave, int& iu, int& iv) const {
    int cv, cu;
    divmod(i, w, rw, cv, cu);
    if (!interleave) {
      iu = u0 + cu;
      iv = v0 + cv;
      return true;
    }
    const int su = (cu >= mu) + (cu >= 2 * mu), sv = (cv >= mv) + (cv >= 2 * mv);
    const int du = 3 * (cu - (int)__umul24((unsigned)su, (unsigned)mu)) + su;
    const int dv = 3 * (cv - (int)__umul24((unsigned)sv, (unsigned)mv)) + sv;
    iu = u0 + du;
    iv = v0 + dv;
    return du < nur && dv < nvr;
  }
};

// (int64_t)v for an integral double |v| < 2^51: v + 1.5 * 2^52 lies in
// [2^52, 2^53), where doubles are spaced by 1, so the sum is exact and its
// encoding is that of 1.5 * 2^52 plus v (one FP64 add and one integer
// subtract instead of the split conversion the compiler emits).
__device__ __forceinline__ int64_t integral_to_i64(double v) {
  const double t = v + 6755399441055744.0;
  return __double_as_longlong(t) - 0x4338000000000000ll;
}

// Exact run [ka, kb) of the ray's samples whose floor index lies in
// [flo, fhi], and the fixed-point position of sample ka relative to the
// tile origin org (in units of 2^-32 voxel) with the fixed-point step.
__device__ __forceinline__ bool ray_run(const Geo& g, const View& v, int iu, int iv,
                                        const int flo[3], const int fhi[3],
                                        const double org[3], int& ka, int& kb, int64_t X[3],
                                        int64_t D[3]) {
  double d[3], P0[3];
  int kend;
  ray_core(g, v, iu, iv, d, P0, kend);
  double Pd[3], Dd[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Pd[c] = rint(P0[c] * kFix);  // == to_fix(P0[c]), kept as an integral double
    Dd[c] = rint(d[c] * kFix);
  }
  exact_window(Pd, Dd, flo, fhi, kend, ka, kb);
  if (kb <= ka) return false;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    // integers below 2^53 throughout, so exact; the results are tile-relative
    // (-2^32 <= X < 2^38 inside the exact window: negative only for the
    // floor -1 samples of partitioned bricks) and |D| <= 2^32
    X[c] = integral_to_i64(Pd[c] + (double)ka * Dd[c] - org[c]);
    D[c] = integral_to_i64(Dd[c]);
  }
  return true;
}

// Stages view a's FP64 vectors into st.sv[slot] (one lane); ordered before
// their readers by the first barrier of the view's sort, and the other slot
// keeps the previously processed view for lanes still in its pass 2 (the
// slot alternates per processed view, so its last readers have passed that
// view's sort barriers).
__device__ __forceinline__ void stage_fp64_view(const Geo& g, int a, int slot,
                                                ChunkStage& st) {
  if (threadIdx.x == 0) {
    View v;
    load_view(g, a, v);
    st.sv[slot] = v;
  }
}

// A over one brick of one volume.  vb: this volume, sbase: its sinogram
// (accumulated into with atomics), tile: the (bx+1)*(by+1)*(bz+1) voxels
// [o, o+b] in LDS (FP32, z stride fwd_tile_stz).
//
// CBX/CBY/CBZ: the brick shape when known at compile time (0 = take the
// run-time value); the instantiated shapes turn every index computation into
// shifts and immediates.
typedef __attribute__((address_space(3))) float lds_f32;

template <int CBX, int CBY, int CBZ, int kR = kSchedR>
__device__ __forceinline__ void fwd_brick_body(const float* __restrict__ vb,
                                               float* __restrict__ sbase, const Geo& g,
                                               int bx_, int by_, int bz_, int nbx, int nby,
                                               int brick, int part, int nparts, float* tile,
                                               ChunkStage& st, RaySched& rs,
                                               long long* __restrict__ fix = nullptr,
                                               double fix_scale = 0.0) {
  const int bx = CBX ? CBX : bx_, by = CBY ? CBY : by_, bz = CBZ ? CBZ : bz_;
  const int jx = brick % nbx, jy = (brick / nbx) % nby, jz = brick / (nbx * nby);
  // bricks tile floor(p) over [-1, N-1]; the tile holds voxels [o, o+b]
  const int ox = -1 + jx * bx, oy = -1 + jy * by, oz = -1 + jz * bz;
  const int ex = min(bx, g.nx - ox), ey = min(by, g.ny - oy), ez = min(bz, g.nz - oz);
  const int tx = bx + 1, ty = by + 1, tz = bz + 1;
  const int sty = fwd_tile_sty(bx), stz = fwd_tile_stz(bx, by);
  __syncthreads();  // the previous item of a persistent workgroup is done with LDS
  for (int i = threadIdx.x; i < tx * ty * tz; i += blockDim.x) {
    const int lx = i % tx, ly = (i / tx) % ty, lz = i / (tx * ty);
    const int x = ox + lx, y = oy + ly, z = oz + lz;
    float val = 0.f;
    if ((unsigned)x < (unsigned)g.nx && (unsigned)y < (unsigned)g.ny &&
        (unsigned)z < (unsigned)g.nz)
      val = vb[((size_t)z * g.ny + y) * g.nx + x];
    tile[lz * stz + ly * sty + lx] = val;
  }
  // samples with floor(p) in [o, o+e-1] belong to this brick
  const int flo[3] = {ox, oy, oz}, fhi[3] = {ox + ex - 1, oy + ey - 1, oz + ez - 1};
  const double lo[3] = {(double)ox, (double)oy, (double)oz};
  const double hi[3] = {(double)(ox + ex), (double)(oy + ey), (double)(oz + ez)};
  const float blo[3] = {ox - kLineSlack, oy - kLineSlack, oz - kLineSlack};
  const float bhi[3] = {ox + ex + kLineSlack, oy + ey + kLineSlack, oz + ez + kLineSlack};
  const double org[3] = {ox * kFix, oy * kFix, oz * kFix};
  const float scale = (float)g.scale;
  const unsigned tbase = (unsigned)(size_t)(lds_f32*)tile;  // LDS byte address

  int nlive = g.n_views;
  if (culls_views(g)) {
    cull_views(g, lo, hi, st);
    nlive = st.n_live;
  }
  int vslot = 0;  // st.sv slot of the next processed view
  const int vend = view_part_end(nlive, part, nparts);
  for (int c0 = view_part_end(nlive, part - 1, nparts); c0 < vend; c0 += kChunk) {
    const int nc = min(kChunk, vend - c0);
    __syncthreads();  // the previous chunk's readers are done with st
    stage_chunk(g, c0, nc, lo, hi, false, st);
    __syncthreads();
    for (int vi = 0; vi < nc; ++vi) {
      const int a = st.vidx[vi];
      RectMap rm;
      rm.init(st.rect[vi], false);
      if (rm.nur <= 0 || rm.nvr <= 0) continue;
      stage_fp64_view(g, a, vslot, st);
      const int cur = vslot;
      vslot ^= 1;
      float fv[12];
#pragma unroll
      for (int k = 0; k < 12; ++k) fv[k] = uni(st.fview[vi][k]);
      float* srow = sbase + (size_t)a * g.nv * g.nu;
      for (int seg = 0; seg < rm.n; seg += kR * blockDim.x) {
        unsigned slot[kR];
#pragma unroll
        for (int j = 0; j < kR; ++j) {
          // one candidate at a time: interleaving the unrolled candidates would
          // only raise register pressure
          asm volatile("" ::: "memory");
          const int i = seg + threadIdx.x + j * blockDim.x;
          slot[j] = ~0u;
          int iu, iv;
          if (i < rm.n && rm.pixel(i, false, iu, iv)) {
            const float ch = chord_samples(fv, iu, iv, blo, bhi);
            if (ch >= 0.f) slot[j] = sched_hit(rs, ch);
          }
        }
        sched_sort(rs, slot);
        const int nhit = uni((int)rs.n);
        View v;
        uni_view(st.sv[cur], v);
        for (int p0 = 0, odd = 0; p0 < nhit; p0 += blockDim.x, odd ^= 1) {
          const int p = pass2_index(p0, odd);
          if (p >= nhit) continue;
          int iu, iv;
          rm.pixel(seg + rs.list[p], false, iu, iv);
          int ka, kb;
          int64_t X[3], D[3];
          if (!ray_run(g, v, iu, iv, flo, fhi, org, ka, kb, X, D)) continue;
          float acc = 0.f;
          for (int k = ka; k < kb; ++k) {
            const int ix = (int)(X[0] >> 32), iy = (int)(X[1] >> 32), iz = (int)(X[2] >> 32);
            const float fx = (float)(uint32_t)X[0] * kInvFix;
            const float fy = (float)(uint32_t)X[1] * kInvFix;
            const float fz = (float)(uint32_t)X[2] * kInvFix;
            // tile indices are small and non-negative (the exact window
            // keeps them in the tile): full-rate 24-bit multiplies
            // LDS byte address of the lower corner: one shift-add and two
            // multiply-adds, the strides in scalar registers (left to
            // itself the compiler re-associates this into six instructions
            // and a constant copy per sample); the upper plane gets its own
            // base so that both planes' pairs use small DS offsets
            unsigned ab;
            asm("v_lshl_add_u32 %0, %1, 2, %2\n\t"
                "v_mad_u32_u24 %0, %3, %4, %0\n\t"
                "v_mad_u32_u24 %0, %5, %6, %0"
                : "=&v"(ab)
                : "v"(ix), "s"(tbase), "v"(iy), "s"((unsigned)(4 * sty)), "v"(iz),
                  "s"((unsigned)(4 * stz)));
            unsigned ab1 = ab + 4u * (unsigned)stz;
            asm volatile("" : "+v"(ab1));
            const lds_f32* t0 = (const lds_f32*)(size_t)ab;
            const lds_f32* t1 = (const lds_f32*)(size_t)ab1;
            const float c000 = t0[0], c100 = t0[1], c010 = t0[sty], c110 = t0[sty + 1];
            const float c001 = t1[0], c101 = t1[1], c011 = t1[sty], c111 = t1[sty + 1];
            const float a0 = c000 + fx * (c100 - c000);
            const float a1 = c010 + fx * (c110 - c010);
            const float a2 = c001 + fx * (c101 - c001);
            const float a3 = c011 + fx * (c111 - c011);
            const float b0 = a0 + fy * (a1 - a0);
            const float b1 = a2 + fy * (a3 - a2);
            acc += b0 + fz * (b1 - b0);
            X[0] += D[0];
            X[1] += D[1];
            X[2] += D[2];
          }
          if (fix)  // integer accumulation: bitwise reproducible
            atomicAdd(reinterpret_cast<unsigned long long*>(
                          fix + (size_t)a * g.nv * g.nu + (size_t)iv * g.nu + iu),
                      (unsigned long long)(long long)llrint((double)acc * fix_scale));
          else
            unsafeAtomicAdd(srow + (size_t)iv * g.nu + iu, acc * scale);
        }
      }
    }
  }
}

// fix != nullptr: the per-(ray, brick) sums go into int64 fixed point at the
// scale fix_ray_scale derives from hdr (max |volume|), unless the volume holds
// a non-finite value, in which case they go into sino as FP32 atomics.
// Workgroups of up to kFwdBlock lanes, at least six waves per SIMD (80
// VGPRs): a tile's LDS lets three workgroups share a CU, so 512-lane ones keep
// twice the waves of 256-lane ones in flight (the spills this costs are all
// outside the sample loop); pass 1 then takes 4 candidates per lane, so one
// segment still fills the schedule list.
constexpr int kFwdBlock = 512;
template <int CBX, int CBY, int CBZ>
__global__ __launch_bounds__(kFwdBlock, 6) void k_fwd_brick(const float* __restrict__ vol,
                                                   float* __restrict__ sino,
                                                   long long* __restrict__ fix,
                                                   const unsigned* __restrict__ hdr, Geo g,
                                                   int bx, int by, int bz, int nbx, int nby) {
  extern __shared__ float tile[];
  __shared__ ChunkStage st;
  __shared__ RaySched rs;
  if (threadIdx.x < kSchedBins) rs.hist[threadIdx.x] = 0u;
  const int b = blockIdx.y;
  double qs = 0.0;
  if (fix) {
    const double S = fix_ray_scale(hdr[0], g);
    if (S > 0.0) {
      qs = g.scale * S;
      fix += (size_t)b * g.n_views * g.nv * g.nu;
    } else {
      fix = nullptr;
    }
  }
  fwd_brick_body<CBX, CBY, CBZ, kSchedCap / kFwdBlock>(
      vol + (size_t)b * g.nx * g.ny * g.nz, sino + (size_t)b * g.n_views * g.nv * g.nu, g, bx,
      by, bz, nbx, nby, blockIdx.x, blockIdx.z, gridDim.z, tile, st, rs, fix, qs);
}

// max |v| over the volume as FP32 bits (non-negative floats order like their
// bits; NaN sorts above infinity), for the fixed-point scale
__global__ void k_absmax(const float* __restrict__ v, size_t n, unsigned* __restrict__ hdr) {
  unsigned m = 0u;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    m = max(m, __float_as_uint(v[i]) & 0x7fffffffu);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, 64));
  if ((threadIdx.x & 63) == 0 && m) atomicMax(hdr, m);
}

// sino = fix / S for a finite volume (the FP32 path wrote sino otherwise)
__global__ void k_fix_to_float(const long long* __restrict__ fix, float* __restrict__ sino,
                               size_t n, const unsigned* __restrict__ hdr, Geo g) {
  const double S = fix_ray_scale(hdr[0], g);
  if (!(S > 0.0)) return;
  const double inv = 1.0 / S;  // a power of two: exact
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    sino[i] = (float)((double)fix[i] * inv);
}


// Nearest integer, halves rounded up (v_cvt_rpi_i32_f32: one instruction
// where the compiler would emit a round and a convert); deterministic, which
// is all the fixed-point accumulation needs.
__device__ __forceinline__ int round_to_int(float x) {
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// 1 + f as an FP32 value for the 32.32 fraction f = lo / 2^32, truncated to
// its top 23 bits: the exponent bits of 1.0 above lo >> 9, one v_alignbit.
__device__ __forceinline__ float one_plus_frac(int64_t x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_alignbit(0x7Fu, (uint32_t)x, 9u));
}

// Adds the trilinear taps of samples [ka, kb) of one ray (value val, tile-
// relative 32.32 position X, step D) into the LDS tile: int32 fixed point
// (kFixed, taps pre-scaled) or FP32.  Each split w -> (w - f w, f w) takes
// f w = fma(1 + f, w, -w), rounded once.
// kGuard (partitioned tiles of bricks on the volume's y = 0 or z = 0 face,
// which have no row or plane below the brick): a sample with floor -1 on y or
// z drops its taps on that outside row or plane and adds the others at
// index 0 (the upper weight moves to the lower slot, the upper slot adds 0).
template <bool kFixed, bool kGuard = false>
__device__ __forceinline__ void adj_march(int ka, int kb, int64_t X[3], const int64_t D[3],
                                          float val, int stz, int sty, int* iacc) {
  float* facc = reinterpret_cast<float*>(iacc);
  for (int k = ka; k < kb; ++k) {
    const int ix = (int)(X[0] >> 32);
    int iy = (int)(X[1] >> 32), iz = (int)(X[2] >> 32);
    const float gx = one_plus_frac(X[0]), gy = one_plus_frac(X[1]), gz = one_plus_frac(X[2]);
    float wz1 = fmaf(gz, val, -val), wz0 = val - wz1;
    if (kGuard && iz < 0) {
      wz0 = wz1;
      wz1 = 0.f;
      iz = 0;
    }
    float w10 = fmaf(gy, wz0, -wz0), w00 = wz0 - w10;
    float w11 = fmaf(gy, wz1, -wz1), w01 = wz1 - w11;
    if (kGuard && iy < 0) {
      w00 = w10;
      w10 = 0.f;
      w01 = w11;
      w11 = 0.f;
      iy = 0;
    }
    const float t001 = fmaf(gx, w00, -w00), t000 = w00 - t001;
    const float t011 = fmaf(gx, w10, -w10), t010 = w10 - t011;
    const float t101 = fmaf(gx, w01, -w01), t100 = w01 - t101;
    const float t111 = fmaf(gx, w11, -w11), t110 = w11 - t111;
    const int base =
        (int)(__umul24((unsigned)iz, (unsigned)stz) + __umul24((unsigned)iy, (unsigned)sty)) + ix;
    if (kFixed) {
      atomicAdd(iacc + base, round_to_int(t000));
      atomicAdd(iacc + base + 1, round_to_int(t001));
      atomicAdd(iacc + base + sty, round_to_int(t010));
      atomicAdd(iacc + base + sty + 1, round_to_int(t011));
      atomicAdd(iacc + base + stz, round_to_int(t100));
      atomicAdd(iacc + base + stz + 1, round_to_int(t101));
      atomicAdd(iacc + base + stz + sty, round_to_int(t110));
      atomicAdd(iacc + base + stz + sty + 1, round_to_int(t111));
    } else {
      atomicAdd(facc + base, t000);
      atomicAdd(facc + base + 1, t001);
      atomicAdd(facc + base + sty, t010);
      atomicAdd(facc + base + sty + 1, t011);
      atomicAdd(facc + base + stz, t100);
      atomicAdd(facc + base + stz + 1, t101);
      atomicAdd(facc + base + stz + sty, t110);
      atomicAdd(facc + base + stz + sty + 1, t111);
    }
    X[0] += D[0];
    X[1] += D[1];
    X[2] += D[2];
  }
}

// Shapes whose owned voxels i = tid + 256 j (j < kAdjPerThread) sit, for
// every lane, at compile-time offsets from one lane base: 256 lanes cover
// whole rows (256 % bx == 0), a step of j whole row groups of a plane
// (by % (256 / bx) == 0), and the 32 slots exactly the brick.
__host__ __device__ constexpr bool fast_flush_shape(int bx, int by, int bz) {
  return bx > 0 && by > 0 && bz > 0 && 256 % bx == 0 && by % (256 / bx) == 0 &&
         bx * by * bz == kAdjPerThread * 256;
}

// Adds the group's owned fixed-point (or FP32) sums into the lane's FP32
// registers and clears them (fixed order per voxel: bitwise reproducible).
// The lane index is made opaque so the addresses are computed here rather
// than hoisted out of the view loops and held (or spilled) throughout.
// kPart: partitioned tile (no column, row or plane below the brick).
template <int CBX, int CBY, int CBZ, bool kPart>
__device__ __forceinline__ void flush_owned(int* iacc, float (&vacc)[kAdjPerThread], bool fixed,
                                            float qinv, int bx, int by, int nown, int stz,
                                            int sty) {
  float* facc = reinterpret_cast<float*>(iacc);
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  constexpr bool kFast = fast_flush_shape(CBX, CBY, CBZ);
  if (kFast && blockDim.x == 256) {
    // voxel j of lane tid: row q + R (j % P) of plane j / P, q = tid / bx
    constexpr int R = kFast ? 256 / CBX : 1, P = kFast ? CBY / R : 1;
    constexpr int SY = kPart ? adj_ptile_sty(CBX) : adj_tile_sty(CBX),
                  SZ = kPart ? adj_ptile_stz(CBX, CBY) : adj_tile_stz(CBX, CBY);
    constexpr int O = kPart ? 0 : SZ + SY + 1;  // tile index of the brick's first voxel
    const int base = O + (tid / (kFast ? CBX : 1)) * SY + tid % (kFast ? CBX : 1);
    if (fixed) {
#pragma unroll
      for (int j = 0; j < kAdjPerThread; ++j) {
        const int li = base + (j % P) * R * SY + (j / P) * SZ;
        vacc[j] += (float)iacc[li] * qinv;
        iacc[li] = 0;
      }
    } else {
#pragma unroll
      for (int j = 0; j < kAdjPerThread; ++j) {
        const int li = base + (j % P) * R * SY + (j / P) * SZ;
        vacc[j] += facc[li];
        iacc[li] = 0;
      }
    }
    return;
  }
  const int o = kPart ? 0 : stz + sty + 1;
#pragma unroll
  for (int j = 0; j < kAdjPerThread; ++j) {
    const int i = tid + j * blockDim.x;
    if (i < nown) {
      const int lx = i % bx, ly = (i / bx) % by, lz = i / (bx * by);
      const int li = lz * stz + ly * sty + lx + o;
      vacc[j] += fixed ? (float)iacc[li] * qinv : facc[li];
      iacc[li] = 0;
    }
  }
}

// Upper-halo voxel h of a brick (0 <= h < adj_halo_count) in local
// coordinates: first the plane lz == bz (x fastest), then the rows ly == by of
// the planes below it, then the column lx == bx of the remaining rows.
__host__ __device__ inline void halo_voxel(int h, int bx, int by, int bz, int& lx, int& ly,
                                           int& lz) {
  const int A = (bx + 1) * (by + 1), B = (bx + 1) * bz;
  if (h < A) {
    lz = bz;
    ly = h / (bx + 1);
    lx = h - ly * (bx + 1);
  } else if (h < A + B) {
    const int r = h - A;
    lz = r / (bx + 1);
    ly = by;
    lx = r - lz * (bx + 1);
  } else {
    const int r = h - A - B;
    lz = r / by;
    ly = r - lz * by;
    lx = bx;
  }
}

// Inverse of halo_voxel for a local voxel with lx == bx, ly == by or lz == bz.
__host__ __device__ inline int halo_slot(int lx, int ly, int lz, int bx, int by, int bz) {
  const int A = (bx + 1) * (by + 1), B = (bx + 1) * bz;
  if (lz == bz) return ly * (bx + 1) + lx;
  if (ly == by) return A + lz * (bx + 1) + lx;
  return A + B + lz * by + ly;
}

// Partitioned tiles: adds the group's upper-halo sums into the FP32
// accumulators hacc (slot h belongs to lane h % blockDim.x throughout, so the
// per-voxel order is fixed) and clears them.
__device__ __forceinline__ void flush_halo(int* iacc, float* hacc, bool fixed, float qinv,
                                           int bx, int by, int bz, int stz, int sty) {
  const float* facc = reinterpret_cast<const float*>(iacc);
  const int nh = adj_halo_count(bx, by, bz);
  for (int h = threadIdx.x; h < nh; h += blockDim.x) {
    int lx, ly, lz;
    halo_voxel(h, bx, by, bz, lx, ly, lz);
    const int li = lz * stz + ly * sty + lx;
    hacc[h] += fixed ? (float)iacc[li] * qinv : facc[li];
    iacc[li] = 0;
  }
}

// Matched A^T over one brick of one volume.  sbase: this volume's sinogram,
// vb: its volume (the brick's voxels are written once), iacc: fixed-point
// int32 accumulators in LDS over the brick plus a one-voxel halo, so every
// tap of every sample in a ray's run is an unconditional atomic (the halo is
// never read back).  Rows are padded to odd strides (bx+3, (bx+3)*(by+2)+1)
// so lanes differing in y or z use different banks.
//
// Views are staged kChunk at a time and accumulated in groups under one
// power-of-two scale: the group's bound on a voxel's sum, sum over its views
// of (weight bound x largest |value| in the view's rectangle), is kept below
// 2^31 / 1.001, so no accumulator can overflow, and a group only grows while
// that sum stays within twice its largest view's bound, so its step is never
// coarser than one view's at 2^30 (the flush traffic shrinks without any
// precision cost; uniform views pair up); after the group the
// owned integers are scaled back into per-lane FP32 registers and cleared.
// Integer addition is order-independent, so the result is bitwise
// reproducible.  A group whose rays carry a non-finite value, or that holds a
// view seeing the box from right next to the source (loose bound),
// accumulates in FP32 LDS atomics in the same buffer instead, so infinities
// and NaNs propagate and precision holds.
//
// kPart (partitioned): the brick marches only the samples whose floor cell is
// its own (floor in [o, o+e-1], or from -1 on the volume's lower faces), so
// every sample is marched by exactly one brick.  Their taps reach the upper
// halo (local coordinate b on some axis), whose sums a group flush adds into
// FP32 accumulators hacc in LDS (after the tile) and which end up in hout for
// the neighbours.  The weight bound staged for the box [floor range, o+e]
// covers those voxels too: they collect only the brick's own samples.  The
// tile has no column, row or plane below the brick: taps of floor -1 samples
// on x land in the previous row's pad word (or the leading one), on y or z
// the guarded loop drops them.
template <int CBX, int CBY, int CBZ, bool kPart = false>
__device__ __forceinline__ void adj_brick_body(const float* __restrict__ sbase,
                                               float* __restrict__ vb, const Geo& g, int bx_,
                                               int by_, int bz_, int nbx, int nby, int brick,
                                               int part, int nparts, bool add, int* iacc,
                                               ChunkStage& st, RaySched& rs,
                                               float* __restrict__ hout = nullptr) {
  const int bx = CBX ? CBX : bx_, by = CBY ? CBY : by_, bz = CBZ ? CBZ : bz_;
  const int jx = brick % nbx, jy = (brick / nbx) % nby, jz = brick / (nbx * nby);
  const int x0 = jx * bx, y0 = jy * by, z0 = jz * bz;
  const int ex = min(bx, g.nx - x0), ey = min(by, g.ny - y0), ez = min(bz, g.nz - z0);
  const int sty = kPart ? adj_ptile_sty(bx) : adj_tile_sty(bx),
            stz = kPart ? adj_ptile_stz(bx, by) : adj_tile_stz(bx, by);
  const int ntile = stz * (kPart ? bz + 1 : bz + 2);
  const int nhalo = kPart ? adj_halo_count(bx, by, bz) : 0;
  int* tile = kPart ? iacc + 1 : iacc;  // partitioned: after the leading pad word
  float* hacc = reinterpret_cast<float*>(tile + ntile);
  const int nown = bx * by * bz;
  __syncthreads();  // the previous item of a persistent workgroup is done with LDS
  for (int i = threadIdx.x; i < (kPart ? 1 : 0) + ntile + nhalo; i += blockDim.x) iacc[i] = 0;
  float vacc[kAdjPerThread];
#pragma unroll
  for (int j = 0; j < kAdjPerThread; ++j) vacc[j] = 0.f;
  // samples with floor(p) in [x0-1, x0+ex-1] touch owned voxels; partitioned,
  // those with floor(p) in [x0, x0+ex-1] (from -1 on the first brick of an axis)
  const int flo[3] = {kPart && jx ? x0 : x0 - 1, kPart && jy ? y0 : y0 - 1,
                      kPart && jz ? z0 : z0 - 1};
  const int fhi[3] = {x0 + ex - 1, y0 + ey - 1, z0 + ez - 1};
  const double lo[3] = {(double)flo[0], (double)flo[1], (double)flo[2]};
  const double hi[3] = {(double)(x0 + ex), (double)(y0 + ey), (double)(z0 + ez)};
  const float blo[3] = {flo[0] - kLineSlack, flo[1] - kLineSlack, flo[2] - kLineSlack};
  const float bhi[3] = {x0 + ex + kLineSlack, y0 + ey + kLineSlack, z0 + ez + kLineSlack};
  // tile origin: the halo corner (x0-1, y0-1, z0-1); partitioned, the
  // brick's first voxel
  const double org[3] = {(kPart ? x0 : x0 - 1) * kFix, (kPart ? y0 : y0 - 1) * kFix,
                         (kPart ? z0 : z0 - 1) * kFix};
  // partitioned bricks on the y = 0 or z = 0 face take the guarded loop
  const bool guard = kPart && (jy == 0 || jz == 0);
  const float gscale = (float)g.scale;

  int nlive = g.n_views;
  if (culls_views(g)) {
    cull_views(g, lo, hi, st);
    nlive = st.n_live;
  }
  int vslot = 0;  // st.sv slot of the next processed view
  const int vend = view_part_end(nlive, part, nparts);
  for (int c0 = view_part_end(nlive, part - 1, nparts); c0 < vend; c0 += kChunk) {
    const int nc = min(kChunk, vend - c0);
    __syncthreads();  // the previous chunk's readers are done with st
    stage_chunk(g, c0, nc, lo, hi, true, st);
    __syncthreads();
    // largest |value| of each view's rectangle (non-finite -> +inf): one wave
    // per view, four loads in flight per lane
    {
      const int lane = threadIdx.x & 63, nw = blockDim.x >> 6;
      for (int vi = threadIdx.x >> 6; vi < nc; vi += nw) {
        const int* rc = st.rect[vi];
        const int nur = rc[1] - rc[0] + 1, nvr = rc[3] - rc[2] + 1;
        if (nur <= 0 || nvr <= 0) continue;
        const float* srow = sbase + (size_t)st.vidx[vi] * g.nv * g.nu;
        const float rn = 1.f / (float)nur;
        const int n = nur * nvr;
        float m = 0.f;
        for (int r0 = lane; r0 < n; r0 += 256) {
          float s[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int r = r0 + 64 * q;
            s[q] = 0.f;
            if (r < n) {
              int dv, du;
              divmod(r, nur, rn, dv, du);
              s[q] = srow[(size_t)(rc[2] + dv) * g.nu + rc[0] + du];
            }
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            m = fmaxf(m, fabsf(s[q] * gscale));
            if (!(fabsf(s[q]) <= 3.0e38f)) m = __int_as_float(0x7f800000);
          }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
        if (lane == 0) st.vmax[vi] = __float_as_uint(m);
      }
    }
    __syncthreads();
    for (int g0 = 0, ng = 0; g0 < nc; g0 += ng) {
      // the group: consecutive views while their summed bound stays within
      // twice the largest one's (its step at 2^31 is then no coarser than a
      // single view's at 2^30); views with no values join for free, a
      // non-finite view or bound ends up alone
      float bound = 0.f, bmax = 0.f;
      bool finite = true;
      ng = 0;
      for (int vi = g0; vi < nc; ++vi) {
        const float vm = __uint_as_float(st.vmax[vi]);
        const float b = vm > 0.f ? st.wbound[vi] * vm : 0.f;
        const bool fin = vm <= 3.0e38f && b <= 3.0e38f;
        const float nb = bound + b, nm = fmaxf(bmax, b);
        if (ng > 0 && (!fin || !finite || nb > 2.f * nm)) break;
        bound = nb;
        bmax = nm;
        finite = finite && fin;
        ++ng;
      }
      const bool fixed = finite && bound <= 3.0e38f;
      float qscale = 1.f, qinv = 1.f;
      if (fixed) {
        // bound * (1 + 1e-3) < 2^e, so sums stay below 2^31 / 1.001: the
        // headroom (> 2e6 units) dwarfs the rounding of the contributions
        // (half a unit each, at most ~5/3 contributions per unit of the
        // weight bound)
        int e;
        frexpf(bound * 1.001f, &e);
        qscale = ldexpf(1.f, 31 - e);
        qinv = ldexpf(1.f, e - 31);
      }
      const float vs = gscale * qscale;
      bool any = false;
      for (int vi = g0; vi < g0 + ng; ++vi) {
        if (st.vmax[vi] == 0u) continue;  // empty rectangle or all-zero values
        any = true;
        const int a = st.vidx[vi];
        RectMap rm;
        rm.init(st.rect[vi], true);
        stage_fp64_view(g, a, vslot, st);
        const int cur = vslot;
        vslot ^= 1;
        float fv[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) fv[k] = uni(st.fview[vi][k]);
        const float* srow = sbase + (size_t)a * g.nv * g.nu;
        for (int seg = 0; seg < rm.n; seg += kSchedR * blockDim.x) {
          unsigned slot[kSchedR];
#pragma unroll
          for (int j = 0; j < kSchedR; ++j) {
            // one candidate at a time: interleaving the unrolled candidates would
            // only raise register pressure
            asm volatile("" ::: "memory");
            const int i = seg + threadIdx.x + j * blockDim.x;
            slot[j] = ~0u;
            int iu, iv;
            if (i < rm.n && rm.pixel(i, true, iu, iv) && srow[(size_t)iv * g.nu + iu] != 0.f) {
              const float ch = chord_samples(fv, iu, iv, blo, bhi);
              if (ch >= 0.f) slot[j] = sched_hit(rs, ch);
            }
          }
          sched_sort(rs, slot);
          const int nhit = uni((int)rs.n);
          View v;
          uni_view(st.sv[cur], v);
          for (int p0 = 0, odd = 0; p0 < nhit; p0 += blockDim.x, odd ^= 1) {
            const int p = pass2_index(p0, odd);
            if (p >= nhit) continue;
            int iu, iv;
            rm.pixel(seg + rs.list[p], true, iu, iv);
            const float val = srow[(size_t)iv * g.nu + iu] * vs;
            int ka, kb;
            int64_t X[3], D[3];
            if (!ray_run(g, v, iu, iv, flo, fhi, org, ka, kb, X, D)) continue;
            // the group's accumulation mode (and the brick's guard) is
            // uniform: one loop per mode, so the sample loop carries no
            // branch on it
            if (kPart && guard) {
              if (fixed)
                adj_march<true, true>(ka, kb, X, D, val, stz, sty, tile);
              else
                adj_march<false, true>(ka, kb, X, D, val, stz, sty, tile);
            } else if (fixed) {
              adj_march<true>(ka, kb, X, D, val, stz, sty, tile);
            } else {
              adj_march<false>(ka, kb, X, D, val, stz, sty, tile);
            }
          }
        }
      }
      if (!any) continue;
      __syncthreads();
      // flush the group's owned voxels into the lane's FP32 registers (fixed
      // group order: the result is bitwise reproducible) and clear them; the
      // halo keeps whatever it gathered and is never read (partitioned: the
      // upper halo goes to its FP32 accumulators likewise).  The lane index is
      // made opaque so the 32 LDS addresses are recomputed here rather than
      // hoisted out of the view loops and held (or spilled) throughout.
      flush_owned<CBX, CBY, CBZ, kPart>(tile, vacc, fixed, qinv, bx, by, nown, stz, sty);
      if (kPart) flush_halo(tile, hacc, fixed, qinv, bx, by, bz, stz, sty);
      __syncthreads();  // cleared before the next group's atomics
    }
  }
#pragma unroll
  for (int j = 0; j < kAdjPerThread; ++j) {
    const int i = threadIdx.x + j * blockDim.x;
    if (i < nown) {
      const int lx = i % bx, ly = (i / bx) % by, lz = i / (bx * by);
      if (lx < ex && ly < ey && lz < ez) {
        float* o = vb + ((size_t)(z0 + lz) * g.ny + (y0 + ly)) * g.nx + (x0 + lx);
        if (!add)
          *o = vacc[j];
        else
          unsafeAtomicAdd(o, vacc[j]);  // the output was zero-filled
      }
    }
  }
  if (kPart) {
    // each lane stores the halo slots it accumulated (no barrier needed)
    for (int h = threadIdx.x; h < nhalo; h += blockDim.x) hout[h] = hacc[h];
  }
}

// At least 3 waves per SIMD (<= 168 VGPRs).  ws == nullptr: the parts (grid
// z) add into vol (zero-filled when there is more than one); else part p
// stores its partial volumes into ws[p][batch][voxels], which k_sum_parts
// then adds up in part order.
template <int CBX, int CBY, int CBZ>
__global__ __launch_bounds__(256, 3) void k_adj_brick(const float* __restrict__ sino,
                                                   float* __restrict__ vol,
                                                   float* __restrict__ ws, Geo g, int bx,
                                                   int by, int bz, int nbx, int nby) {
  extern __shared__ int iacc[];
  __shared__ ChunkStage st;
  __shared__ RaySched rs;
  if (threadIdx.x < kSchedBins) rs.hist[threadIdx.x] = 0u;
  const int b = blockIdx.y;
  const size_t nvox = (size_t)g.nx * g.ny * g.nz;
  float* out = ws ? ws + ((size_t)blockIdx.z * gridDim.y + b) * nvox : vol + (size_t)b * nvox;
  adj_brick_body<CBX, CBY, CBZ>(sino + (size_t)b * g.n_views * g.nv * g.nu, out, g, bx, by, bz,
                                nbx, nby, blockIdx.x, blockIdx.z, gridDim.z,
                                !ws && gridDim.z > 1, iacc, st, rs);
}

// Partitioned A^T (adj_brick_body with kPart).  Owned sums: with `into`
// set, every part stores (one part) or adds (two, into the zero-filled
// output: order-independent) into vol; else part 0 stores into vol and part
// p >= 1 into wsv[p - 1][batch][voxels].  Every part stores its upper-halo
// sums into halo[part][batch][brick][adj_halo_count].
template <int CBX, int CBY, int CBZ>
__global__ __launch_bounds__(256, 3) void k_adj_part(const float* __restrict__ sino,
                                                  float* __restrict__ vol,
                                                  float* __restrict__ wsv,
                                                  float* __restrict__ halo, Geo g, int bx,
                                                  int by, int bz, int nbx, int nby, int into) {
  extern __shared__ int iacc[];
  __shared__ ChunkStage st;
  __shared__ RaySched rs;
  if (threadIdx.x < kSchedBins) rs.hist[threadIdx.x] = 0u;
  const int b = blockIdx.y, p = blockIdx.z;
  const size_t nvox = (size_t)g.nx * g.ny * g.nz;
  float* out = into || !p ? vol + (size_t)b * nvox : wsv + ((size_t)(p - 1) * gridDim.y + b) * nvox;
  const size_t nh = adj_halo_count(CBX ? CBX : bx, CBY ? CBY : by, CBZ ? CBZ : bz);
  float* hout = halo + (((size_t)p * gridDim.y + b) * gridDim.x + blockIdx.x) * nh;
  adj_brick_body<CBX, CBY, CBZ, true>(sino + (size_t)b * g.n_views * g.nv * g.nu, out, g, bx,
                                      by, bz, nbx, nby, blockIdx.x, p, gridDim.z,
                                      into && gridDim.z > 1, iacc, st, rs, hout);
}

// Adds to s the upper-halo sums that the lower neighbours of brick (jx, jy,
// jz) hold for its voxel (lx, ly, lz) (fx, fy, fz: the voxel is on the
// brick's lower x, y, z face and that neighbour exists): offsets d = 1..7
// (bit 0 x, bit 1 y, bit 2 z), each summed over the parts in part order, added
// in the order of d (one fixed order per voxel: bitwise reproducible).  All
// loads of a part are issued before they are added.
__device__ __forceinline__ float add_halos(float s, const float* __restrict__ halo, int b,
                                           int batch, int parts, int jx, int jy, int jz, int lx,
                                           int ly, int lz, bool fx, bool fy, bool fz, int bx,
                                           int by, int bz, int nbx, int nby, int nbz) {
  const size_t nh = adj_halo_count(bx, by, bz);
  const size_t nb = (size_t)nbx * nby * nbz;
  size_t off[7];
  bool ok[7];
  float t[7];
#pragma unroll
  for (int d = 1; d < 8; ++d) {
    const int dx = d & 1, dy = (d >> 1) & 1, dz = d >> 2;
    ok[d - 1] = (!dx || fx) && (!dy || fy) && (!dz || fz);
    const size_t src = ((size_t)(jz - dz) * nby + (jy - dy)) * nbx + (jx - dx);
    off[d - 1] =
        ok[d - 1] ? src * nh + halo_slot(lx + dx * bx, ly + dy * by, lz + dz * bz, bx, by, bz)
                  : 0;
    t[d - 1] = 0.f;
  }
  for (int p = 0; p < parts; ++p) {
    const float* hp = halo + ((size_t)p * batch + b) * nb * nh;
    float h[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) h[d] = ok[d] ? hp[off[d]] : 0.f;
#pragma unroll
    for (int d = 0; d < 7; ++d) t[d] += h[d];
  }
#pragma unroll
  for (int d = 0; d < 7; ++d)
    if (ok[d]) s += t[d];
  return s;
}

// Completes the partitioned A^T in place when vol holds part 0's owned sums
// and wsv the other parts': every voxel gets the parts in part order, then
// add_halos.  One lane per voxel of an x row, blockIdx.y = y, blockIdx.z =
// z + nz (b - b0).
__global__ void k_sum_halo(float* __restrict__ vol, const float* __restrict__ wsv,
                           const float* __restrict__ halo, Geo g, int bx, int by, int bz,
                           int nbx, int nby, int nbz, int parts, int batch, int b0) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z % g.nz, b = b0 + (int)blockIdx.z / g.nz;
  if (x >= g.nx) return;
  const int jx = x / bx, jy = y / by, jz = z / bz;
  const int lx = x - jx * bx, ly = y - jy * by, lz = z - jz * bz;
  const bool fx = lx == 0 && jx > 0, fy = ly == 0 && jy > 0, fz = lz == 0 && jz > 0;
  const size_t nvox = (size_t)g.nx * g.ny * g.nz;
  const size_t v = (size_t)b * nvox + ((size_t)z * g.ny + y) * g.nx + x;
  float s = vol[v];
  for (int p = 1; p < parts; ++p) s += wsv[(size_t)(p - 1) * batch * nvox + v];
  if (fx || fy || fz)
    s = add_halos(s, halo, b, batch, parts, jx, jy, jz, lx, ly, lz, fx, fy, fz, bx, by, bz, nbx,
                  nby, nbz);
  vol[v] = s;
}

// Completes the partitioned A^T in place when vol already holds every part's
// owned sums: only the voxels on a brick's lower faces change, one lane each
// (the z faces, then the y faces off the z faces, then the x faces off both;
// blockIdx.y = batch entry).
__global__ void k_halo_faces(float* __restrict__ vol, const float* __restrict__ halo, Geo g,
                             int bx, int by, int bz, int nbx, int nby, int nbz, int parts,
                             int batch) {
  const int b = blockIdx.y;
  const long long nzf = (long long)(nbz - 1) * g.ny * g.nx;
  const long long nyf = (long long)(nby - 1) * g.nz * g.nx;
  const long long nxf = (long long)(nbx - 1) * g.nz * g.ny;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int x, y, z;
  if (i < nzf) {
    const int pi = (int)(i / g.nx), q = (int)(i - (long long)pi * g.nx);
    x = q;
    y = pi % g.ny;
    z = (pi / g.ny + 1) * bz;
  } else if ((i -= nzf) < nyf) {
    const int pi = (int)(i / g.nx), q = (int)(i - (long long)pi * g.nx);
    x = q;
    z = pi % g.nz;
    y = (pi / g.nz + 1) * by;
    if (z % bz == 0 && z > 0) return;  // on a z face
  } else if ((i -= nyf) < nxf) {
    const int pi = (int)(i / g.ny), q = (int)(i - (long long)pi * g.ny);
    y = q;
    z = pi % g.nz;
    x = (pi / g.nz + 1) * bx;
    if ((z % bz == 0 && z > 0) || (y % by == 0 && y > 0)) return;  // on a z or y face
  } else {
    return;
  }
  const int jx = x / bx, jy = y / by, jz = z / bz;
  const int lx = x - jx * bx, ly = y - jy * by, lz = z - jz * bz;
  const bool fx = lx == 0 && jx > 0, fy = ly == 0 && jy > 0, fz = lz == 0 && jz > 0;
  const size_t v = (size_t)b * g.nx * g.ny * g.nz + ((size_t)z * g.ny + y) * g.nx + x;
  vol[v] = add_halos(vol[v], halo, b, batch, parts, jx, jy, jz, lx, ly, lz, fx, fy, fz, bx, by,
                     bz, nbx, nby, nbz);
}

// out[i] = ws[0][i] + ws[1][i] + ... in part order (FP32, one fixed order: the
// split adjoint is bitwise reproducible for any number of parts)
__global__ void k_sum_parts(const float* __restrict__ ws, float* __restrict__ out, size_t n,
                            int parts) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    float s = ws[i];
    for (int p = 1; p < parts; ++p) s += ws[(size_t)p * n + i];
    out[i] = s;
  }
}

// ---- uneven instances in one launch ----------------------------------------

__device__ Geo instance_geo(const dtv_instance& in) {
  Geo g{};
  g.nx = in.nx;
  g.ny = in.ny;
  g.nz = in.nz;
  g.nu = in.nu;
  g.nv = in.nv;
  g.n_views = in.n_views;
  g.axis_mode = in.axis_mode;
  for (int c = 0; c < 3; ++c) {
    g.lo[c] = in.box_lo[c];
    g.hi[c] = in.box_hi[c];
  }
  g.margin = in.margin;
  g.scale = in.scale;
  g.kind = kTable;
  g.table = in.views;
  return g;
}

// Next work item: blockIdx.x in one-item-per-workgroup mode, else taken
// from the device counter (persistent workgroups, dynamic balance).
__device__ __forceinline__ int next_item(int* counter, int* slot, int done) {
  if (!counter) return done == 0 ? (int)blockIdx.x : 0x7fffffff;
  __syncthreads();
  if (threadIdx.x == 0) *slot = atomicAdd(counter, 1);
  __syncthreads();
  return *slot;
}

// Workgroups of up to kFwdBlock lanes, as k_fwd_brick.
template <int CBX, int CBY, int CBZ>
__global__ __launch_bounds__(kFwdBlock, 6) void k_fwd_brick_multi(const dtv_instance* __restrict__ inst,
                                                         const dtv_work_item* __restrict__ work,
                                                         int n_work, int* counter) {
  extern __shared__ float tile[];
  __shared__ ChunkStage st;
  __shared__ RaySched rs;
  __shared__ int slot;
  if (threadIdx.x < kSchedBins) rs.hist[threadIdx.x] = 0u;
  for (int done = 0;; ++done) {
    const int w = next_item(counter, &slot, done);
    if (w >= n_work) break;
    const dtv_work_item it = work[w];
    const dtv_instance in = inst[it.instance];
    const Geo g = instance_geo(in);
    const int nb = in.nbx * in.nby * in.nbz;
    fwd_brick_body<CBX, CBY, CBZ, kSchedCap / kFwdBlock>(
        in.in, in.out, g, in.bx, in.by, in.bz, in.nbx, in.nby, it.brick % nb, it.brick / nb,
        max(1, in.view_parts), tile, st, rs);
  }
}

template <int CBX, int CBY, int CBZ>
__global__ __launch_bounds__(256, 3) void k_adj_brick_multi(const dtv_instance* __restrict__ inst,
                                                         const dtv_work_item* __restrict__ work,
                                                         int n_work, int* counter) {
  extern __shared__ int iacc[];
  __shared__ ChunkStage st;
  __shared__ RaySched rs;
  __shared__ int slot;
  if (threadIdx.x < kSchedBins) rs.hist[threadIdx.x] = 0u;
  for (int done = 0;; ++done) {
    const int w = next_item(counter, &slot, done);
    if (w >= n_work) break;
    const dtv_work_item it = work[w];
    const dtv_instance in = inst[it.instance];
    const Geo g = instance_geo(in);
    const int nb = in.nbx * in.nby * in.nbz;
    adj_brick_body<CBX, CBY, CBZ>(in.in, in.out, g, in.bx, in.by, in.bz, in.nbx, in.nby,
                                  it.brick % nb, it.brick / nb, max(1, in.view_parts),
                                  in.view_parts > 1, iacc, st, rs);
  }
}

// Partitioned matched A^T over uneven instances (adj_brick_body with kPart):
// the owned sums go into the instance's volume (stored, or added by two parts
// into zeros), the upper-halo sums of item (brick, part) to halo +
// hoff[instance] + (part x bricks + brick) x adj_halo_count.
template <int CBX, int CBY, int CBZ>
__global__ __launch_bounds__(256, 3) void k_adj_part_multi(const dtv_instance* __restrict__ inst,
                                                        const dtv_work_item* __restrict__ work,
                                                        int n_work, int* counter,
                                                        const long long* __restrict__ hoff,
                                                        float* __restrict__ halo) {
  extern __shared__ int iacc[];
  __shared__ ChunkStage st;
  __shared__ RaySched rs;
  __shared__ int slot;
  if (threadIdx.x < kSchedBins) rs.hist[threadIdx.x] = 0u;
  for (int done = 0;; ++done) {
    const int w = next_item(counter, &slot, done);
    if (w >= n_work) break;
    const dtv_work_item it = work[w];
    const dtv_instance in = inst[it.instance];
    const Geo g = instance_geo(in);
    const int nb = in.nbx * in.nby * in.nbz;
    const size_t nh = adj_halo_count(CBX ? CBX : in.bx, CBY ? CBY : in.by, CBZ ? CBZ : in.bz);
    float* hout = halo + hoff[it.instance] + (size_t)it.brick * nh;
    adj_brick_body<CBX, CBY, CBZ, true>(in.in, in.out, g, in.bx, in.by, in.bz, in.nbx, in.nby,
                                        it.brick % nb, it.brick / nb, max(1, in.view_parts),
                                        in.view_parts > 1, iacc, st, rs, hout);
  }
}

// Lanes on a brick's lower faces of an instance (k_halo_faces' enumeration).
__host__ __device__ inline long long face_lanes(int nx, int ny, int nz, int nbx, int nby,
                                                int nbz) {
  return (long long)(nbz - 1) * ny * nx + (long long)(nby - 1) * nz * nx +
         (long long)(nbx - 1) * nz * ny;
}

// Workspace header of the partitioned multi-instance A^T, from the device
// table (so a captured graph recomputes it): hoff[i] = floats of halo sums
// before instance i (view parts x bricks x adj_halo_count each), foff[i] =
// face-kernel workgroups before instance i (whole workgroups per instance).
// One workgroup: each lane sums a contiguous run of instances, lane 0 scans.
__global__ void k_multi_offsets(const dtv_instance* __restrict__ inst, int n,
                                long long* __restrict__ hoff, long long* __restrict__ foff) {
  __shared__ long long th[256], tf[256];
  const int t = threadIdx.x, per = (n + 255) / 256;
  const int a = min(n, t * per), e = min(n, a + per);
  long long sh = 0, sf = 0;
  for (int i = a; i < e; ++i) {
    const dtv_instance in = inst[i];
    sh += (long long)max(1, in.view_parts) * in.nbx * in.nby * in.nbz *
          adj_halo_count(in.bx, in.by, in.bz);
    sf += (face_lanes(in.nx, in.ny, in.nz, in.nbx, in.nby, in.nbz) + 255) / 256;
  }
  th[t] = sh;
  tf[t] = sf;
  __syncthreads();
  if (t == 0) {
    long long ch = 0, cf = 0;
    for (int k = 0; k < 256; ++k) {
      const long long h = th[k], f = tf[k];
      th[k] = ch;
      tf[k] = cf;
      ch += h;
      cf += f;
    }
    hoff[n] = ch;
    foff[n] = cf;
  }
  __syncthreads();
  sh = th[t];
  sf = tf[t];
  for (int i = a; i < e; ++i) {
    const dtv_instance in = inst[i];
    hoff[i] = sh;
    foff[i] = sf;
    sh += (long long)max(1, in.view_parts) * in.nbx * in.nby * in.nbz *
          adj_halo_count(in.bx, in.by, in.bz);
    sf += (face_lanes(in.nx, in.ny, in.nz, in.nbx, in.nby, in.nbz) + 255) / 256;
  }
}

// k_halo_faces over every instance of a multi launch: workgroup w belongs to
// the instance i with foff[i] <= w < foff[i + 1] (found by bisection); its
// lanes enumerate that instance's face voxels as k_halo_faces does.
__global__ void k_halo_faces_multi(const dtv_instance* __restrict__ inst, int n,
                                   const long long* __restrict__ hoff,
                                   const long long* __restrict__ foff,
                                   const float* __restrict__ halo) {
  const long long w = blockIdx.x;
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (foff[mid] <= w) lo = mid;
    else hi = mid;
  }
  const dtv_instance in = inst[lo];
  const int bx = in.bx, by = in.by, bz = in.bz, nx = in.nx, ny = in.ny, nz = in.nz;
  const int nbx = in.nbx, nby = in.nby, nbz = in.nbz;
  const long long nzf = (long long)(nbz - 1) * ny * nx, nyf = (long long)(nby - 1) * nz * nx,
                  nxf = (long long)(nbx - 1) * nz * ny;
  long long i = (w - foff[lo]) * blockDim.x + threadIdx.x;
  int x, y, z;
  if (i < nzf) {
    const int pi = (int)(i / nx), q = (int)(i - (long long)pi * nx);
    x = q;
    y = pi % ny;
    z = (pi / ny + 1) * bz;
  } else if ((i -= nzf) < nyf) {
    const int pi = (int)(i / nx), q = (int)(i - (long long)pi * nx);
    x = q;
    z = pi % nz;
    y = (pi / nz + 1) * by;
    if (z % bz == 0 && z > 0) return;
  } else if ((i -= nyf) < nxf) {
    const int pi = (int)(i / ny), q = (int)(i - (long long)pi * ny);
    y = q;
    z = pi % nz;
    x = (pi / nz + 1) * bx;
    if ((z % bz == 0 && z > 0) || (y % by == 0 && y > 0)) return;
  } else {
    return;
  }
  const int jx = x / bx, jy = y / by, jz = z / bz;
  const int lx = x - jx * bx, ly = y - jy * by, lz = z - jz * bz;
  const bool fx = lx == 0 && jx > 0, fy = ly == 0 && jy > 0, fz = lz == 0 && jz > 0;
  float* o = in.out + ((size_t)z * ny + y) * nx + x;
  *o = add_halos(*o, halo + hoff[lo], 0, 1, max(1, in.view_parts), jx, jy, jz, lx, ly, lz, fx,
                 fy, fz, bx, by, bz, nbx, nby, nbz);
}

// Adjoint instances whose views are split: out = 0 (the parts add into it),
// blockIdx.y = instance.
__global__ void k_zero_split_multi(const dtv_instance* __restrict__ inst) {
  const dtv_instance in = inst[blockIdx.y];
  if (in.view_parts <= 1) return;
  const size_t n = (size_t)in.nx * in.ny * in.nz;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    in.out[i] = 0.f;
}

// out[i] = -init[i] (or 0) for every instance's output, blockIdx.y = instance
__global__ void k_init_multi(const dtv_instance* __restrict__ inst) {
  const dtv_instance in = inst[blockIdx.y];
  const size_t n = (size_t)in.n_views * in.nu * in.nv;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    in.out[i] = in.init ? -in.init[i] : 0.f;
}

__global__ void k_zero_int(int* p) { *p = 0; }

}  // namespace

namespace dtv {

int launch_fwd_brick(const float* vol, float* sino, int batch, const Geo& g, int bx, int by,
                     int bz, int block, int parts, hipStream_t st, void* fix_ws) {
  if (block > kFwdBlock) return 1;
  const int nbx = (g.nx + 1 + bx - 1) / bx, nby = (g.ny + 1 + by - 1) / by,
            nbz = (g.nz + 1 + bz - 1) / bz;
  const size_t lds = fwd_tile_bytes(bx, by, bz);
  auto kern = &k_fwd_brick<0, 0, 0>;
#define DTV_PICK(a, b, c) \
  if (bx == a && by == b && bz == c) kern = &k_fwd_brick<a, b, c>;
  DTV_BRICK_SHAPES(DTV_PICK)
#undef DTV_PICK
  // fix_ws: 256-byte header (max |vol| bits), then one int64 per ray; the
  // caller zero-filled both and the sinogram
  unsigned* hdr = static_cast<unsigned*>(fix_ws);
  long long* fix = fix_ws ? reinterpret_cast<long long*>(static_cast<char*>(fix_ws) + 256)
                          : nullptr;
  if (fix_ws) {
    const size_t nv = (size_t)batch * g.nx * g.ny * g.nz;
    const unsigned blocks = (unsigned)std::min<size_t>((nv + 255) / 256, 4096);
    hipLaunchKernelGGL(k_absmax, dim3(blocks), dim3(256), 0, st, vol, nv, hdr);
  }
  hipLaunchKernelGGL(kern, dim3(nbx * nby * nbz, batch, parts), dim3(block), lds, st, vol, sino,
                     fix, hdr, g, bx, by, bz, nbx, nby);
  if (fix_ws) {
    const size_t ns = (size_t)batch * g.n_views * g.nu * g.nv;
    const unsigned blocks = (unsigned)std::min<size_t>((ns + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fix_to_float, dim3(blocks), dim3(256), 0, st, fix, sino, ns, hdr, g);
  }
  return hipGetLastError() == hipSuccess ? 0 : 4;
}

int launch_adj_brick(const float* sino, float* vol, float* ws, int batch, const Geo& g, int bx,
                     int by, int bz, int block, int parts, hipStream_t st) {
  const int nbx = (g.nx + bx - 1) / bx, nby = (g.ny + by - 1) / by, nbz = (g.nz + bz - 1) / bz;
  if (bx * by * bz > kAdjPerThread * block) return 1;
  const size_t lds = adj_tile_bytes(bx, by, bz);
  auto kern = &k_adj_brick<0, 0, 0>;
#define DTV_PICK(a, b, c) \
  if (bx == a && by == b && bz == c) kern = &k_adj_brick<a, b, c>;
  DTV_BRICK_SHAPES(DTV_PICK)
#undef DTV_PICK
  hipLaunchKernelGGL(kern, dim3(nbx * nby * nbz, batch, parts), dim3(block), lds, st, sino, vol,
                     ws, g, bx, by, bz, nbx, nby);
  if (hipGetLastError() != hipSuccess) return 4;
  if (ws && parts > 1) {
    const size_t n = (size_t)batch * g.nx * g.ny * g.nz;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_sum_parts, dim3(blocks), dim3(256), 0, st, ws, vol, n, parts);
  }
  return hipGetLastError() == hipSuccess ? 0 : 4;
}

size_t adj_part_ws_bytes(int batch, const Geo& g, int bx, int by, int bz, int parts) {
  const size_t nvox = (size_t)g.nx * g.ny * g.nz;
  const size_t nb = (size_t)((g.nx + bx - 1) / bx) * ((g.ny + by - 1) / by) * ((g.nz + bz - 1) / bz);
  return ((size_t)parts * batch * nb * adj_halo_count(bx, by, bz) +
          (size_t)(parts - 1) * batch * nvox) *
         sizeof(float);
}

int launch_adj_part(const float* sino, float* vol, float* ws, int batch, const Geo& g, int bx,
                    int by, int bz, int block, int parts, bool into, hipStream_t st) {
  const int nbx = (g.nx + bx - 1) / bx, nby = (g.ny + by - 1) / by, nbz = (g.nz + bz - 1) / bz;
  if (bx * by * bz > kAdjPerThread * block || parts < 1 || (into && parts > 2)) return 1;
  const size_t lds = adj_ptile_bytes(bx, by, bz);
  if (lds > kMaxTileBytes) return 1;
  auto kern = &k_adj_part<0, 0, 0>;
#define DTV_PICK(a, b, c) \
  if (bx == a && by == b && bz == c) kern = &k_adj_part<a, b, c>;
  DTV_BRICK_SHAPES(DTV_PICK)
#undef DTV_PICK
  const size_t nb = (size_t)nbx * nby * nbz;
  float* halo = ws;
  float* wsv = ws + (size_t)parts * batch * nb * adj_halo_count(bx, by, bz);
  hipLaunchKernelGGL(kern, dim3(nbx * nby * nbz, batch, parts), dim3(block), lds, st, sino, vol,
                     wsv, halo, g, bx, by, bz, nbx, nby, into ? 1 : 0);
  if (hipGetLastError() != hipSuccess) return 4;
  if (!into && parts > 1) {
    const int tb = std::min(256, (g.nx + 63) / 64 * 64);
    const int per = std::max(1, 65535 / g.nz);  // batch entries per launch (grid z <= 65535)
    for (int b0 = 0; b0 < batch; b0 += per) {
      const int n = std::min(per, batch - b0);
      hipLaunchKernelGGL(k_sum_halo, dim3((g.nx + tb - 1) / tb, g.ny, g.nz * n), dim3(tb), 0, st,
                         vol, wsv, halo, g, bx, by, bz, nbx, nby, nbz, parts, batch, b0);
      if (hipGetLastError() != hipSuccess) return 4;
    }
  } else {
    const long long nf = (long long)(nbz - 1) * g.ny * g.nx + (long long)(nby - 1) * g.nz * g.nx +
                         (long long)(nbx - 1) * g.nz * g.ny;
    if (nf > 0) {
      hipLaunchKernelGGL(k_halo_faces, dim3((unsigned)((nf + 255) / 256), batch), dim3(256), 0,
                         st, vol, halo, g, bx, by, bz, nbx, nby, nbz, parts, batch);
      if (hipGetLastError() != hipSuccess) return 4;
    }
  }
  return 0;
}

}  // namespace dtv

namespace {

int validate_multi(const dtv_instance* ih, int32_t n_inst, const dtv_work_item* wh,
                   int32_t n_work, int32_t block, bool adjoint, size_t& lds,
                   bool part = false) {
  if (!ih || !wh) return DTV_ERR_NULL_POINTER;
  if (n_inst < 1 || n_work < 1 || n_inst > 65535) return DTV_ERR_INVALID_ARGUMENT;
  // A's multi kernel takes workgroups of up to 512 lanes, A^T's up to 256
  if (block < 64 || block > (adjoint ? 256 : kFwdBlock) || block % 64)
    return DTV_ERR_INVALID_ARGUMENT;
  lds = 0;
  size_t max_out = 0;
  for (int i = 0; i < n_inst; ++i) {
    const dtv_instance& in = ih[i];
    if (!in.in || !in.out || !in.views) return DTV_ERR_NULL_POINTER;
    if (in.nx < 1 || in.ny < 1 || in.nz < 1 || in.nu < 1 || in.nv < 1 || in.n_views < 1)
      return DTV_ERR_INVALID_ARGUMENT;
    if ((int64_t)in.nx * in.ny * in.nz >= (1ll << 31) || (int64_t)in.nu * in.nv >= (1ll << 31))
      return DTV_ERR_TOO_LARGE;
    if (in.axis_mode != kAxisXY && in.axis_mode != kAxisXYZ) return DTV_ERR_INVALID_ARGUMENT;
    if (in.bx < 2 || in.by < 2 || in.bz < 2 || in.bx > 64 || in.by > 64 || in.bz > 64)
      return DTV_ERR_INVALID_ARGUMENT;
    const int ext = adjoint ? 0 : 1;
    if (in.nbx != (in.nx + ext + in.bx - 1) / in.bx || in.nby != (in.ny + ext + in.by - 1) / in.by ||
        in.nbz != (in.nz + ext + in.bz - 1) / in.bz)
      return DTV_ERR_INVALID_ARGUMENT;
    size_t l;
    if (adjoint) {
      if (in.bx * in.by * in.bz > kAdjPerThread * block) return DTV_ERR_INVALID_ARGUMENT;
      l = part ? adj_ptile_bytes(in.bx, in.by, in.bz) : adj_tile_bytes(in.bx, in.by, in.bz);
    } else {
      l = fwd_tile_bytes(in.bx, in.by, in.bz);
    }
    if (l > kMaxTileBytes) return DTV_ERR_INVALID_ARGUMENT;
    if (in.view_parts < 0 || in.view_parts > (adjoint ? 2 : 16)) return DTV_ERR_INVALID_ARGUMENT;
    lds = l > lds ? l : lds;
    const size_t no = (size_t)in.n_views * in.nu * in.nv;
    max_out = no > max_out ? no : max_out;
  }
  for (int w = 0; w < n_work; ++w) {
    if (wh[w].instance < 0 || wh[w].instance >= n_inst) return DTV_ERR_INVALID_ARGUMENT;
    const dtv_instance& in = ih[wh[w].instance];
    const long long units = (long long)in.nbx * in.nby * in.nbz * std::max(1, in.view_parts);
    if (wh[w].brick < 0 || wh[w].brick >= units) return DTV_ERR_INVALID_ARGUMENT;
  }
  (void)max_out;
  return DTV_OK;
}

// Workgroups of a multi launch: one per item, or for the persistent queue
// persistent_blocks capped at what the device holds at once for this kernel
// (workgroups beyond that would only start once the queue is empty).
unsigned multi_grid(const void* kern, int block, size_t lds, bool persistent,
                    int32_t persistent_blocks, int32_t n_work) {
  if (!persistent) return (unsigned)n_work;
  int n = std::min(persistent_blocks, n_work);
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, lds) == hipSuccess &&
      cus > 0 && per_cu > 0)
    n = std::min(n, cus * per_cu);
  return (unsigned)(n > 0 ? n : 1);
}

// Workspace of the partitioned multi-instance A^T: a header of hoff[n + 1]
// and foff[n + 1] (int64, k_multi_offsets) rounded up to 256 bytes, then the
// halo sums.
size_t multi_ws_header(int32_t n_inst) { return ((size_t)16 * (n_inst + 1) + 255) / 256 * 256; }
size_t multi_ws_bytes(const dtv_instance* ih, int32_t n_inst) {
  size_t h = 0;
  for (int i = 0; i < n_inst; ++i)
    h += (size_t)std::max(1, ih[i].view_parts) * ih[i].nbx * ih[i].nby * ih[i].nbz *
         adj_halo_count(ih[i].bx, ih[i].by, ih[i].bz);
  return multi_ws_header(n_inst) + h * sizeof(float);
}

int launch_multi(bool adjoint, const dtv_instance* ih, const dtv_instance* id, int32_t n_inst,
                 const dtv_work_item* wh, const dtv_work_item* wd, int32_t n_work, int32_t block,
                 int32_t* counter, int32_t persistent_blocks, hipStream_t st,
                 void* ws = nullptr, size_t ws_bytes = 0) {
  const bool part = adjoint && ws;
  size_t lds = 0;
  int rc = validate_multi(ih, n_inst, wh, n_work, block, adjoint, lds, part);
  if (rc) return rc;
  if (part) {
    if ((uintptr_t)ws % 256 || ws_bytes < multi_ws_bytes(ih, n_inst))
      return DTV_ERR_INVALID_ARGUMENT;
    if (lds > kMaxTileBytes) return DTV_ERR_INVALID_ARGUMENT;
  }
  if (!id || !wd) return DTV_ERR_NULL_POINTER;
  if (!adjoint) {
    for (int i = 0; i < n_inst; ++i) {
      rc = check_atomic_target(ih[i].out);
      if (rc) return rc;
    }
    size_t max_out = 0;
    for (int i = 0; i < n_inst; ++i) {
      const size_t no = (size_t)ih[i].n_views * ih[i].nu * ih[i].nv;
      max_out = no > max_out ? no : max_out;
    }
    const unsigned gx = (unsigned)std::min<size_t>((max_out + 255) / 256, 4096);
    hipLaunchKernelGGL(k_init_multi, dim3(gx, n_inst), dim3(256), 0, st, id);
    if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
  } else {
    size_t max_split = 0;
    for (int i = 0; i < n_inst; ++i) {
      if (ih[i].view_parts <= 1) continue;
      rc = check_atomic_target(ih[i].out);
      if (rc) return rc;
      const size_t nv = (size_t)ih[i].nx * ih[i].ny * ih[i].nz;
      max_split = nv > max_split ? nv : max_split;
    }
    if (max_split) {
      const unsigned gx = (unsigned)std::min<size_t>((max_split + 255) / 256, 4096);
      hipLaunchKernelGGL(k_zero_split_multi, dim3(gx, n_inst), dim3(256), 0, st, id);
      if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
    }
  }
  const bool persistent = counter && persistent_blocks > 0;
  if (persistent) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, counter);
    if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
  }
  int* ctr = persistent ? counter : nullptr;
  // one compiled shape when every instance uses the same listed brick shape
  bool same = true;
  for (int i = 1; i < n_inst; ++i)
    same = same && ih[i].bx == ih[0].bx && ih[i].by == ih[0].by && ih[i].bz == ih[0].bz;
  const int bx = ih[0].bx, by = ih[0].by, bz = ih[0].bz;
  if (part) {
    long long* hoff = static_cast<long long*>(ws);
    long long* foff = hoff + (n_inst + 1);
    float* halo = reinterpret_cast<float*>(static_cast<char*>(ws) + multi_ws_header(n_inst));
    hipLaunchKernelGGL(k_multi_offsets, dim3(1), dim3(256), 0, st, id, n_inst, hoff, foff);
    if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
    auto kern = &k_adj_part_multi<0, 0, 0>;
#define DTV_PICK(a, b, c) \
    if (same && bx == a && by == b && bz == c) kern = &k_adj_part_multi<a, b, c>;
    DTV_BRICK_SHAPES(DTV_PICK)
#undef DTV_PICK
    const dim3 grid(multi_grid(reinterpret_cast<const void*>(kern), block, lds, persistent,
                               persistent_blocks, n_work));
    hipLaunchKernelGGL(kern, grid, dim3(block), lds, st, id, wd, n_work, ctr, hoff, halo);
    if (hipGetLastError() != hipSuccess) return DTV_ERR_HIP;
    long long groups = 0;
    for (int i = 0; i < n_inst; ++i)
      groups += (face_lanes(ih[i].nx, ih[i].ny, ih[i].nz, ih[i].nbx, ih[i].nby, ih[i].nbz) + 255) /
                256;
    if (groups > 0x7fffffffll) return DTV_ERR_TOO_LARGE;
    if (groups > 0)
      hipLaunchKernelGGL(k_halo_faces_multi, dim3((unsigned)groups), dim3(256), 0, st, id, n_inst,
                         hoff, foff, halo);
  } else if (adjoint) {
    auto kern = &k_adj_brick_multi<0, 0, 0>;
#define DTV_PICK(a, b, c) \
    if (same && bx == a && by == b && bz == c) kern = &k_adj_brick_multi<a, b, c>;
    DTV_BRICK_SHAPES(DTV_PICK)
#undef DTV_PICK
    const dim3 grid(multi_grid(reinterpret_cast<const void*>(kern), block, lds, persistent,
                               persistent_blocks, n_work));
    hipLaunchKernelGGL(kern, grid, dim3(block), lds, st, id, wd, n_work, ctr);
  } else {
    auto kern = &k_fwd_brick_multi<0, 0, 0>;
#define DTV_PICK(a, b, c) \
    if (same && bx == a && by == b && bz == c) kern = &k_fwd_brick_multi<a, b, c>;
    DTV_BRICK_SHAPES(DTV_PICK)
#undef DTV_PICK
    const dim3 grid(multi_grid(reinterpret_cast<const void*>(kern), block, lds, persistent,
                               persistent_blocks, n_work));
    hipLaunchKernelGGL(kern, grid, dim3(block), lds, st, id, wd, n_work, ctr);
  }
  return hipGetLastError() == hipSuccess ? DTV_OK : DTV_ERR_HIP;
}

}  // namespace

extern "C" {

int dtv_cone_forward_multi(const dtv_instance* inst_host, const dtv_instance* inst_dev,
                           int32_t n_inst, const dtv_work_item* work_host,
                           const dtv_work_item* work_dev, int32_t n_work, int32_t block,
                           int32_t* counter, int32_t persistent_blocks, void* stream) {
  return launch_multi(false, inst_host, inst_dev, n_inst, work_host, work_dev, n_work, block,
                      counter, persistent_blocks, (hipStream_t)stream);
}

int dtv_cone_adjoint_multi(const dtv_instance* inst_host, const dtv_instance* inst_dev,
                           int32_t n_inst, const dtv_work_item* work_host,
                           const dtv_work_item* work_dev, int32_t n_work, int32_t block,
                           int32_t* counter, int32_t persistent_blocks, void* stream) {
  return launch_multi(true, inst_host, inst_dev, n_inst, work_host, work_dev, n_work, block,
                      counter, persistent_blocks, (hipStream_t)stream);
}

int dtv_cone_adjoint_multi_ws(const dtv_instance* inst_host, const dtv_instance* inst_dev,
                              int32_t n_inst, const dtv_work_item* work_host,
                              const dtv_work_item* work_dev, int32_t n_work, int32_t block,
                              int32_t* counter, int32_t persistent_blocks, void* workspace,
                              size_t workspace_bytes, void* stream) {
  if (!workspace) return DTV_ERR_NULL_POINTER;
  return launch_multi(true, inst_host, inst_dev, n_inst, work_host, work_dev, n_work, block,
                      counter, persistent_blocks, (hipStream_t)stream, workspace,
                      workspace_bytes);
}

int dtv_cone_adjoint_multi_ws_bytes(const dtv_instance* inst_host, int32_t n_inst,
                                    size_t* bytes) {
  if (!inst_host || !bytes) return DTV_ERR_NULL_POINTER;
  if (n_inst < 1 || n_inst > 65535) return DTV_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < n_inst; ++i) {
    const dtv_instance& in = inst_host[i];
    if (in.bx < 2 || in.by < 2 || in.bz < 2 || in.bx > 64 || in.by > 64 || in.bz > 64 ||
        in.nbx < 1 || in.nby < 1 || in.nbz < 1 || in.view_parts < 0 || in.view_parts > 2)
      return DTV_ERR_INVALID_ARGUMENT;
  }
  *bytes = multi_ws_bytes(inst_host, n_inst);
  return DTV_OK;
}

}  // extern "C"
