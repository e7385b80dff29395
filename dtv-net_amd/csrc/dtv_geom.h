// Cone-beam geometry shared by the host entry points and the gfx950 kernels.
//
// Every geometry is reduced to per-view FP64 vectors in the voxel frame (the
// frame where voxel (ix, iy, iz) has its sample centre at the integer point
// (ix, iy, iz)): source S, pixel (0,0) position D0, per-pixel steps U and V.
// Rays are normalised; samples sit one voxel apart at
// beta_k = beta_start + k, beta_start = min(alpha0, alpha1) - margin, where
// alpha0/1 are the distances to the box planes of the ray's dominant axis.
// This reproduces the sample phase of
//   Standard/plug/kernel/forwardKernel.cu:30-80 (axis among x,y; ties -> x)
//   Beijing/plug/kernel/forwardKernel.cu:24-73  (axis among x,y,z; ties x->y->z)
// Per-ray set-up runs in FP64 with FP contraction off so that it follows the
// same rounding as the host-side table builder; sample positions are then
// advanced in 32.32 fixed point, which keeps them exact to ~1e-7 voxel over
// any ray length (FP32 stepping would drift by ~1e-5 voxel at 256^3).
#pragma once

#include <stdint.h>
#include <math.h>

#ifndef DTV_HD
#if defined(__HIPCC__)
#define DTV_HD __host__ __device__
#else
#define DTV_HD
#endif
#endif

namespace dtv {

// ``#define PI 3.14159265359`` of the reference kernels (forwardKernel.cu:14).
constexpr double kPiRef = 3.14159265359;
constexpr double kFix = 4294967296.0;            // 2^32
constexpr float kInvFix = 2.3283064365386963e-10f; // 2^-32

enum ViewKind : int { kTable = 0, kStandard = 1, kBeijing = 2 };
enum AxisMode : int { kAxisXY = 2, kAxisXYZ = 3 };

struct View {
  double S[3], D0[3], U[3], V[3];
};

// Launch-time description of one projector call (passed by value).
struct Geo {
  int nx, ny, nz;       // volume size, x fastest
  int nu, nv;           // detector size, u fastest
  int n_views;
  int axis_mode;        // kAxisXY or kAxisXYZ
  int kind;             // ViewKind: where the per-view vectors come from
  double lo[3], hi[3];  // box whose dominant-axis planes phase the samples
  double margin;        // voxels added at both ends of the sample range
  double scale;         // output multiplier
  double sid, sdd;      // kStandard only
  const double* table;  // kTable: [n_views][12] = S, D0, U, V
  const double* angles; // kStandard: degrees
  const float* pvec;    // kBeijing: reference projection vectors [n_views][12]
};

// Standard circular orbit (forwardKernel.cu:30-40, world unit = one pixel,
// voxel = SID/SDD world units, p = world / (SID/SDD) + N/2).
DTV_HD inline void view_standard(double angle, double sid, double sdd, int nx, int ny,
                                 int nz, int nu, int nv, View& v) {
#pragma clang fp contract(off)
  const double s = sid / sdd;
  const double th = angle / 180.0 * kPiRef - kPiRef / 2.0;
  const double c = cos(th), sn = sin(th);
  v.S[0] = -c * sid / s + nx / 2.0;
  v.S[1] = -sn * sid / s + ny / 2.0;
  v.S[2] = nz / 2.0;
  v.D0[0] = v.S[0] + (c * sdd + sn * (nu / 2.0)) / s;
  v.D0[1] = v.S[1] + (sn * sdd - c * (nu / 2.0)) / s;
  v.D0[2] = v.S[2] + (-nv / 2.0) / s;
  v.U[0] = -sn / s;
  v.U[1] = c / s;
  v.U[2] = 0.0;
  v.V[0] = 0.0;
  v.V[1] = 0.0;
  v.V[2] = 1.0 / s;
}

// Beijing projection vector (forwardKernel.cu:27-33): source (0..2) and
// detector centre (3..5) with z negated, v (6..8), u (9..11), pixel centres
// at +0.5, voxel frame p = kernel frame + N/2.
DTV_HD inline void view_beijing(const float* p, int nx, int ny, int nz, int nu, int nv,
                                View& v) {
#pragma clang fp contract(off)
  const double hx = nx / 2.0, hy = ny / 2.0, hz = nz / 2.0;
  v.S[0] = (double)p[0] + hx;
  v.S[1] = (double)p[1] + hy;
  v.S[2] = -(double)p[2] + hz;
  const double dc0 = (double)p[3] + hx;
  const double dc1 = (double)p[4] + hy;
  const double dc2 = -(double)p[5] + hz;
  const double au = 0.5 - nu / 2.0, av = 0.5 - nv / 2.0;
  for (int c = 0; c < 3; ++c) {
    v.U[c] = (double)p[9 + c];
    v.V[c] = (double)p[6 + c];
  }
  v.D0[0] = dc0 + au * v.U[0] + av * v.V[0];
  v.D0[1] = dc1 + au * v.U[1] + av * v.V[1];
  v.D0[2] = dc2 + au * v.U[2] + av * v.V[2];
}

DTV_HD inline void load_view(const Geo& g, int a, View& v) {
  if (g.kind == kStandard) {
    view_standard(g.angles[a], g.sid, g.sdd, g.nx, g.ny, g.nz, g.nu, g.nv, v);
  } else if (g.kind == kBeijing) {
    view_beijing(g.pvec + 12 * (size_t)a, g.nx, g.ny, g.nz, g.nu, g.nv, v);
  } else {
    const double* t = g.table + 12 * (size_t)a;
    for (int c = 0; c < 3; ++c) {
      v.S[c] = t[c];
      v.D0[c] = t[3 + c];
      v.U[c] = t[6 + c];
      v.V[c] = t[9 + c];
    }
  }
}

// One ray, ready to march: position of sample k is P + k*D in 32.32 fixed
// point; samples k in [k0, k1) may touch the volume, every other sample of
// the reference's range lies wholly outside the trilinear support.
struct Ray {
  int64_t P[3];
  int64_t D[3];
  int k0, k1;
};

// 1/x and 1/sqrt(x) to about one ulp.  On the device: the hardware estimate
// refined by two Newton steps in FMA form (each step squares the relative
// error, the final FMA adds at most 2^-53); on the host: the correctly
// rounded operations.
DTV_HD inline double recip(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
#else
  return 1.0 / x;
#endif
}

DTV_HD inline double rsqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);  // 1 - x y^2
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
#else
  return 1.0 / sqrt(x);
#endif
}

// Per-ray set-up shared by every kernel, so that A and the matched A^T see
// bitwise-identical sample positions (the FMAs are written out, so every
// inlined copy rounds the same way).  Outputs the unit direction d, the
// position P0 of sample k = 0 (both voxel frame, FP64) and kend: the
// reference's samples are k in [0, kend) (beta_start + k < beta_end).
DTV_HD inline void ray_core(const Geo& g, const View& v, int iu, int iv, double d[3],
                            double P0[3], int& kend) {
#pragma clang fp contract(off)
  double r[3];
  for (int c = 0; c < 3; ++c)
    r[c] = fma((double)iv, v.V[c], fma((double)iu, v.U[c], v.D0[c])) - v.S[c];
  const double il = rsqrt(fma(r[0], r[0], fma(r[1], r[1], r[2] * r[2])));
  d[0] = r[0] * il;
  d[1] = r[1] * il;
  d[2] = r[2] * il;
  const double adx = fabs(d[0]), ady = fabs(d[1]), adz = fabs(d[2]);
  int ax;
  if (g.axis_mode == kAxisXY) {
    ax = (adx >= ady) ? 0 : 1;
  } else {
    const double m = fmax(adx, fmax(adz, ady));
    ax = (adx == m) ? 0 : ((ady == m) ? 1 : 2);
  }
  // select without dynamic indexing so the view stays in registers
  const double dax = ax == 0 ? d[0] : (ax == 1 ? d[1] : d[2]);
  const double sax = ax == 0 ? v.S[0] : (ax == 1 ? v.S[1] : v.S[2]);
  const double lax = ax == 0 ? g.lo[0] : (ax == 1 ? g.lo[1] : g.lo[2]);
  const double hax = ax == 0 ? g.hi[0] : (ax == 1 ? g.hi[1] : g.hi[2]);
  const double idax = recip(dax);
  const double a0 = (lax - sax) * idax;
  const double a1 = (hax - sax) * idax;
  const double bs = fmin(a0, a1) - g.margin;
  const double be = fmax(a0, a1) + g.margin;
  // samples k >= 0 with bs + k < be; ceil(be - bs) is within one of the
  // exact count, non-finite or absurd ranges give no samples
  kend = 0;
  if (be - bs > 0.0 && be - bs < 1.0e7) {
    kend = (int)ceil(be - bs);
    if (kend > 0 && !(bs + (double)(kend - 1) < be)) --kend;
    if (bs + (double)kend < be) ++kend;
  }
  for (int c = 0; c < 3; ++c) P0[c] = fma(bs, d[c], v.S[c]);
}

DTV_HD inline int64_t to_fix(double x) { return (int64_t)llrint(x * kFix); }

// The samples k in [0, kend) of the fixed-point ray P + k D (32.32) whose
// floor index lies in [lo_c, hi_c] on every axis c form one run [ka, kb)
// (each coordinate is monotone in k); this finds it exactly from the fixed-
// point values themselves, Pd = rint(P0 * 2^32) and Dd = rint(d * 2^32)
// (integral doubles, equal to to_fix(P0), to_fix(d)), so it agrees with the
// sample loop even where the real ray and the fixed-point one fall on
// opposite sides of a voxel plane (a direction component below 2^-33 rounds
// to a zero step).  Every numerator below is an integer under 2^53 and
// |Dd| <= 2^32, so a non-integral quotient lies at least 2^-32 from every
// integer, while n * (1/Dd) carries a relative error under 2^-52 (recip is
// within ~2^-53, the product adds one rounding): for
// quotients under 2^18 in magnitude ceil(t - 2^-33) and floor(t + 2^-33) are
// the exact integer bounds.  The bounds are solved for k - kr; a run that
// starts beyond 2^16 is solved again around its estimated start, and bounds
// far from the run only ever lose to the binding ones.
DTV_HD inline void exact_window(const double Pd[3], const double Dd[3], const int lo[3],
                                const int hi[3], int kend, int& ka, int& kb) {
#pragma clang fp contract(off)
  constexpr double kBias = 1.1641532182693481e-10;  // 2^-33
  double rD[3];
  for (int c = 0; c < 3; ++c) rD[c] = recip(Dd[c]);  // relative error <= ~2^-53
  int kr = 0;
  for (int pass = 0; pass < 2; ++pass) {
    double a = (double)(0 - kr), b = (double)(kend - 1 - kr);
    for (int c = 0; c < 3; ++c) {
      const double pr = pass == 0 ? Pd[c] : Pd[c] + (double)kr * Dd[c];
      const double nl = (double)lo[c] * kFix - pr;               // k D >= nl
      const double nh = ((double)hi[c] + 1.0) * kFix - 1.0 - pr;  // k D <= nh
      if (Dd[c] == 0.0) {
        if (nl > 0.0 || nh < 0.0) b = a - 1.0;
      } else {
        const double r = rD[c];
        const double t0 = nl * r, t1 = nh * r;
        const double tl = Dd[c] > 0.0 ? t0 : t1, th = Dd[c] > 0.0 ? t1 : t0;
        a = fmax(a, ceil(tl - kBias));
        b = fmin(b, floor(th + kBias));
      }
    }
    if (!(b >= a)) break;
    if (a < 65536.0 || pass == 1) {
      ka = kr + (int)a;
      kb = kr + (int)b + 1;
      return;
    }
    kr += (int)a - 1;
  }
  ka = kb = 0;
}

// Whole-ray set-up: fixed-point start and step plus the window [k0, k1)
// of samples whose trilinear support can touch the volume (p in [-1, N) on
// every axis, widened by one sample at each end; every sample is still
// bounds-checked tap by tap where this window is used).
DTV_HD inline void ray_setup(const Geo& g, const View& v, int iu, int iv, Ray& ray) {
#pragma clang fp contract(off)
  double d[3], P0[3];
  int kend;
  ray_core(g, v, iu, iv, d, P0, kend);
  const double n3[3] = {(double)g.nx, (double)g.ny, (double)g.nz};
  double klo = 0.0, khi = (double)kend;
  for (int c = 0; c < 3; ++c) {
    if (d[c] != 0.0) {
      const double t0 = (-1.0 - P0[c]) / d[c];
      const double t1 = (n3[c] - P0[c]) / d[c];
      klo = fmax(klo, fmin(t0, t1) - 1.0);
      khi = fmin(khi, fmax(t0, t1) + 2.0);
    } else if (!(P0[c] >= -1.0 && P0[c] < n3[c])) {
      khi = -1.0;
    }
  }
  if (!(khi > klo)) {
    ray.k0 = ray.k1 = 0;
  } else {
    ray.k0 = (int)floor(klo);
    ray.k1 = (int)ceil(khi);
    if (ray.k1 > kend) ray.k1 = kend;
    if (ray.k0 < 0) ray.k0 = 0;
    if (ray.k1 < ray.k0) ray.k1 = ray.k0;
  }
  for (int c = 0; c < 3; ++c) {
    ray.P[c] = to_fix(P0[c]);
    ray.D[c] = to_fix(d[c]);
  }
}

// Detector-plane projection helpers of one view: a point c maps to the
// continuous pixel coordinates (a, b) with S + t (c - S) = D0 + a U + b V:
// t = h / ((c - S).n), a = t ((c - S).ea) - qa, b = t ((c - S).eb) - qb.
struct ViewProj {
  double n[3], ea[3], eb[3];
  double h, qa, qb;
};

DTV_HD inline void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

DTV_HD inline double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

DTV_HD inline void view_proj(const View& v, ViewProj& p) {
  double vn[3], nu[3], ds[3];
  cross3(v.U, v.V, p.n);
  cross3(v.V, p.n, vn);
  cross3(p.n, v.U, nu);
  const double sa = 1.0 / dot3(v.U, vn), sb = 1.0 / dot3(v.V, nu);
  for (int c = 0; c < 3; ++c) {
    p.ea[c] = vn[c] * sa;
    p.eb[c] = nu[c] * sb;
    ds[c] = v.D0[c] - v.S[c];
  }
  p.h = dot3(ds, p.n);
  p.qa = dot3(ds, p.ea);
  p.qb = dot3(ds, p.eb);
}

// Projects point c; returns false when c is not strictly in front of the
// source (then the caller must take the whole detector).
DTV_HD inline bool project_point(const View& v, const ViewProj& p, const double c[3], double& a,
                                 double& b) {
  const double w[3] = {c[0] - v.S[0], c[1] - v.S[1], c[2] - v.S[2]};
  const double wn = dot3(w, p.n);
  if (!(wn * p.h > 0.0)) return false;
  const double t = p.h / wn;
  a = t * dot3(w, p.ea) - p.qa;
  b = t * dot3(w, p.eb) - p.qb;
  return true;
}

// Bound on sum_k w over the samples of one view that touch one voxel c of a
// box (w = the sample's trilinear weight on c):
//  * per ray, the weights on c are bounded by a 1-D tent along the ray's
//    dominant axis sampled at spacing |d_ax| >= 1/sqrt(3): sum <= 3;
//  * rays touching c pass through its support cube (side 2), so their pixels
//    lie in the cube's projected footprint.  Along u that footprint is at
//    most sum_k (2 / e_k) du_k times the depth ratio across the box
//    (perspective), +10%, where du_k is the largest projected length of the
//    box's edges along axis k and e_k the box's extent along k (triangle
//    inequality over the cube's three edge directions); likewise along v.  A
//    footprint of f pixels holds <= f + 1 pixel centres per axis (+0.5 slack).
// Depths are |(corner - S).n|, proportional to the distance from the source
// along the detector normal.  nur x nvr is the view's candidate rectangle.
DTV_HD inline float weight_per_voxel_bound(bool all, const double du[3], const double dv[3],
                                           double wmin, double wmax, const float ext[3],
                                           int nur, int nvr) {
  if (!all || !(wmin > 0.0)) return 3.0f * (float)nur * (float)nvr;
  const double persp = wmax / wmin * 1.1;
  double fu = 0.0, fv = 0.0;
  for (int k = 0; k < 3; ++k) {
    fu += du[k] * 2.0 / ext[k];
    fv += dv[k] * 2.0 / ext[k];
  }
  const float pu = (float)(fu * persp) + 1.5f, pv = (float)(fv * persp) + 1.5f;
  return 3.0f * fminf(pu, (float)nur + 1.f) * fminf(pv, (float)nvr + 1.f);
}

// What the brick kernels stage per (box, view), from the projections of the
// box [lo, hi]'s 8 corners (corner t at (t&1 ? hi : lo, t&2 ? hi : lo,
// t&4 ? hi : lo) per axis), reduced over the corners: amin..bmax the range
// of their continuous pixel coordinates (project_point), wmin/wmax of their
// depths |(corner - S).n|, du[k]/dv[k] the largest projected length of the
// box's edges along axis k (corner t to corner t ^ (1 << k)), ok = every
// corner strictly in front of the source.
//  * rect = {u0, u1, v0, v1} (inclusive, u1 < u0 when empty): the pixels i
//    in [floor(amin), ceil(amax)] (likewise v), clipped to the detector --
//    pixel i's ray passes through coordinate i exactly, so these are the
//    rays that can meet the box, and floor/ceil give a pixel of slack
//    against the FP64 projection's rounding (~1e-10 pixel); the whole
//    detector when some corner is not in front of the source.
//  * wbound: weight_per_voxel_bound, or +inf when a corner is behind the
//    source or the depth ratio across the box exceeds 2 (the box next to the
//    source, where the bound would be loose): the matched adjoint then
//    accumulates that view in FP32 instead of fixed point.
// The kernels reduce over the corners with wave shuffles (one corner per
// lane) and call this; box_view_stage reduces with loops and calls it.
DTV_HD inline void box_view_finish(double amin, double amax, double bmin, double bmax,
                                   double wmin, double wmax, const double du[3],
                                   const double dv[3], bool ok, const double lo[3],
                                   const double hi[3], int nu, int nv, int rect[4],
                                   float& wbound) {
  int u0 = 0, u1 = nu - 1, v0 = 0, v1 = nv - 1;
  if (ok) {
    const double fu0 = floor(amin), fu1 = ceil(amax);
    const double fv0 = floor(bmin), fv1 = ceil(bmax);
    u0 = fu0 < 0.0 ? 0 : (fu0 > nu ? nu : (int)fu0);
    u1 = fu1 < -1.0 ? -1 : (fu1 > nu - 1 ? nu - 1 : (int)fu1);
    v0 = fv0 < 0.0 ? 0 : (fv0 > nv ? nv : (int)fv0);
    v1 = fv1 < -1.0 ? -1 : (fv1 > nv - 1 ? nv - 1 : (int)fv1);
  }
  rect[0] = u0;
  rect[1] = u1;
  rect[2] = v0;
  rect[3] = v1;
  const float ext[3] = {(float)(hi[0] - lo[0]), (float)(hi[1] - lo[1]), (float)(hi[2] - lo[2])};
  wbound = (ok && wmin > 0.0 && wmax <= 2.0 * wmin)
               ? weight_per_voxel_bound(true, du, dv, wmin, wmax, ext, u1 - u0 + 1, v1 - v0 + 1)
               : INFINITY;
}

// box_view_finish from the 8 corners' values (ca, cb: pixel coordinates,
// cw: depths), reduced with loops.
DTV_HD inline void box_view_stage(const double ca[8], const double cb[8], const double cw[8],
                                  bool ok, const double lo[3], const double hi[3], int nu,
                                  int nv, int rect[4], float& wbound) {
  double amin = ca[0], amax = ca[0], bmin = cb[0], bmax = cb[0], wmin = cw[0], wmax = cw[0];
  double du[3] = {0.0, 0.0, 0.0}, dv[3] = {0.0, 0.0, 0.0};
  for (int t = 0; t < 8; ++t) {
    amin = fmin(amin, ca[t]);
    amax = fmax(amax, ca[t]);
    bmin = fmin(bmin, cb[t]);
    bmax = fmax(bmax, cb[t]);
    wmin = fmin(wmin, cw[t]);
    wmax = fmax(wmax, cw[t]);
    for (int k = 0; k < 3; ++k) {
      du[k] = fmax(du[k], fabs(ca[t] - ca[t ^ (1 << k)]));
      dv[k] = fmax(dv[k], fabs(cb[t] - cb[t ^ (1 << k)]));
    }
  }
  box_view_finish(amin, amax, bmin, bmax, wmin, wmax, du, dv, ok, lo, hi, nu, nv, rect, wbound);
}

}  // namespace dtv
