// Host check of the matched adjoint's fixed-point overflow bound and of the
// staged candidate rectangle (dtv::box_view_stage in
// dtv-net_amd/csrc/dtv_geom.h, the code the brick kernels run per (brick,
// view)).
//
// For a brick of owned voxels [o, o + e) the kernel stages the box
// [o - 1, o + e] (every sample whose floor index lies in [o - 1, o + e - 1]
// touches an owned voxel) and takes from box_view_stage:
//   * rect: the pixels whose rays can reach the box;
//   * wbound: a bound on sum_k w_k(c) over the view's samples, for every
//     owned voxel c (w = trilinear weight), or +inf (then the kernel
//     accumulates the view in FP32).
// The int32 accumulators cannot overflow only if wbound really bounds that
// sum.  This program enumerates every ray of the detector with the kernels'
// own set-up (ray_core, 32.32 fixed-point positions, exact_window), sums the
// weights per owned voxel in FP64 and checks, for every (brick, view):
//   max_c sum_k w_k(c) <= wbound        (when wbound is finite)
//   every ray with a sample in the box lies inside rect.
// Geometries: circular orbits at magnification 1.1 to 10, tilted Beijing-
// style detectors with a source off the mid-plane, and sources next to the
// volume (depth ratios up to 2 across a brick, corners behind the source);
// bricks from 8^3 to 32x32x8 at seeded positions, edge bricks included.
// The partitioned adjoint (each sample marched only by the brick holding its
// floor cell: floor in [o, o + e - 1], from -1 on the first brick of an axis)
// stages the box [o, o + e] (o - 1 on a first brick) and reads back its
// upper halo too, so there the bound must hold for every tile voxel in
// [o, o + e] inside the volume, over the samples of that window; each brick is
// checked both ways.
// Prints "ok <checks> <inf_views> <max_ratio>" or the first failure (exit 1).
// Built and run by tests/test_bound_native.py.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "dtv_geom.h"

using namespace dtv;

struct Case {
  Geo g;
  std::vector<View> views;
};

static Geo base_geo(int nx, int ny, int nz, int nu, int nv) {
  Geo g{};
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.nu = nu;
  g.nv = nv;
  g.axis_mode = kAxisXYZ;
  g.kind = kTable;
  const double n[3] = {(double)nx, (double)ny, (double)nz};
  for (int c = 0; c < 3; ++c) {
    g.lo[c] = -0.5;
    g.hi[c] = n[c] - 0.5;
  }
  g.margin = 3.0;
  g.scale = 1.0;
  return g;
}

// Orbit around the volume centre: source at distance R (voxels), detector
// centre Rd beyond the axis, tilted by (tilt_u, tilt_v) radians, source and
// detector shifted by zoff along z; pixel pitch chosen so the detector
// covers the volume's shadow with 10% margin.
static Case orbit(int nx, int ny, int nz, int nu, int nv, int n_views, double R, double Rd,
                  double tilt_u, double tilt_v, double zoff, double phase) {
  Case cs;
  cs.g = base_geo(nx, ny, nz, nu, nv);
  const double off[3] = {nx / 2.0 - 0.5, ny / 2.0 - 0.5, nz / 2.0 - 0.5};
  const double mag = (R + Rd) / R;
  const double diag = std::sqrt((double)nx * nx + (double)ny * ny);
  const double pu = 1.1 * diag * mag / nu, pv = 1.1 * (nz + 2.0 * std::fabs(zoff)) * mag / nv;
  for (int a = 0; a < n_views; ++a) {
    const double th = phase + 2.0 * M_PI * a / n_views;
    const double ex[3] = {std::cos(th), std::sin(th), 0.0};
    const double ey[3] = {-std::sin(th), std::cos(th), 0.0};
    const double ez[3] = {0.0, 0.0, 1.0};
    // tilt: u rotated towards ex by tilt_u, v rotated towards ex by tilt_v
    double U[3], V[3];
    for (int c = 0; c < 3; ++c) {
      U[c] = (std::cos(tilt_u) * ey[c] + std::sin(tilt_u) * ex[c]) * pu;
      V[c] = (std::cos(tilt_v) * ez[c] + std::sin(tilt_v) * ex[c]) * pv;
    }
    View v;
    for (int c = 0; c < 3; ++c) {
      v.S[c] = -ex[c] * R + off[c] + (c == 2 ? zoff : 0.0);
      const double dc = ex[c] * Rd + off[c] + (c == 2 ? zoff : 0.0);
      v.U[c] = U[c];
      v.V[c] = V[c];
      v.D0[c] = dc + (0.5 - nu / 2.0) * U[c] + (0.5 - nv / 2.0) * V[c];
    }
    cs.views.push_back(v);
  }
  return cs;
}

static int fails = 0;
static long checks = 0, inf_views = 0;
static double max_ratio = 0.0;

static void check_brick(const Case& cs, int bx, int by, int bz, int jx, int jy, int jz,
                        bool part) {
  const Geo& g = cs.g;
  const int x0 = jx * bx, y0 = jy * by, z0 = jz * bz;
  if (x0 >= g.nx || y0 >= g.ny || z0 >= g.nz) return;
  const int ex = std::min(bx, g.nx - x0), ey = std::min(by, g.ny - y0),
            ez = std::min(bz, g.nz - z0);
  const int flo[3] = {part && jx ? x0 : x0 - 1, part && jy ? y0 : y0 - 1,
                      part && jz ? z0 : z0 - 1};
  const int fhi[3] = {x0 + ex - 1, y0 + ey - 1, z0 + ez - 1};
  const double lo[3] = {(double)flo[0], (double)flo[1], (double)flo[2]};
  const double hi[3] = {(double)(x0 + ex), (double)(y0 + ey), (double)(z0 + ez)};
  // voxels whose sums are read back: owned, plus the upper halo inside the
  // volume when partitioned
  const int ax = part ? std::min(ex + 1, g.nx - x0) : ex, ay = part ? std::min(ey + 1, g.ny - y0) : ey,
            az = part ? std::min(ez + 1, g.nz - z0) : ez;
  std::vector<double> acc((size_t)ax * ay * az);
  for (const View& v : cs.views) {
    ViewProj vp;
    view_proj(v, vp);
    double ca[8], cb[8], cw[8];
    bool ok = true;
    for (int t = 0; t < 8; ++t) {
      const double c[3] = {(t & 1) ? hi[0] : lo[0], (t & 2) ? hi[1] : lo[1],
                           (t & 4) ? hi[2] : lo[2]};
      ca[t] = cb[t] = 0.0;
      ok = project_point(v, vp, c, ca[t], cb[t]) && ok;
      const double w[3] = {c[0] - v.S[0], c[1] - v.S[1], c[2] - v.S[2]};
      cw[t] = std::fabs(dot3(w, vp.n));
    }
    int rect[4];
    float wb;
    box_view_stage(ca, cb, cw, ok, lo, hi, g.nu, g.nv, rect, wb);
    std::fill(acc.begin(), acc.end(), 0.0);
    for (int iv = 0; iv < g.nv; ++iv) {
      for (int iu = 0; iu < g.nu; ++iu) {
        double d[3], P0[3];
        int kend;
        ray_core(g, v, iu, iv, d, P0, kend);
        double Pd[3], Dd[3];
        int64_t P[3], D[3];
        for (int c = 0; c < 3; ++c) {
          Pd[c] = std::rint(P0[c] * kFix);
          Dd[c] = std::rint(d[c] * kFix);
          P[c] = (int64_t)Pd[c];
          D[c] = (int64_t)Dd[c];
        }
        int ka = 0, kb = 0;
        exact_window(Pd, Dd, flo, fhi, kend, ka, kb);
        if (kb <= ka) continue;
        if (iu < rect[0] || iu > rect[1] || iv < rect[2] || iv > rect[3]) {
          std::printf("FAIL rect: ray (%d,%d) reaches brick (%d,%d,%d)+(%d,%d,%d) outside "
                      "[%d,%d]x[%d,%d]\n", iu, iv, x0, y0, z0, ex, ey, ez, rect[0], rect[1],
                      rect[2], rect[3]);
          ++fails;
          return;
        }
        for (int k = ka; k < kb; ++k) {
          int64_t f[3];
          double fr[3];
          for (int c = 0; c < 3; ++c) {
            const int64_t X = P[c] + (int64_t)k * D[c];
            f[c] = X >> 32;
            fr[c] = (double)(X & 0xffffffffll) / kFix;
          }
          for (int t = 0; t < 8; ++t) {
            const int64_t ix = f[0] + (t & 1), iy = f[1] + ((t >> 1) & 1),
                          iz = f[2] + ((t >> 2) & 1);
            if (ix < x0 || ix >= x0 + ax || iy < y0 || iy >= y0 + ay || iz < z0 ||
                iz >= z0 + az)
              continue;
            const double w = ((t & 1) ? fr[0] : 1.0 - fr[0]) *
                             (((t >> 1) & 1) ? fr[1] : 1.0 - fr[1]) *
                             (((t >> 2) & 1) ? fr[2] : 1.0 - fr[2]);
            acc[((size_t)(iz - z0) * ay + (iy - y0)) * ax + (ix - x0)] += w;
          }
        }
      }
    }
    double m = 0.0;
    for (double s : acc) m = std::max(m, s);
    if (!std::isfinite(wb)) {
      ++inf_views;
      continue;
    }
    ++checks;
    if (m > 0.0) max_ratio = std::max(max_ratio, m / (double)wb);
    if (m > (double)wb) {
      std::printf("FAIL bound%s: brick (%d,%d,%d)+(%d,%d,%d): max sum w %.6f > bound %.6f\n",
                  part ? " (partitioned)" : "", x0, y0, z0, ex, ey, ez, m, (double)wb);
      ++fails;
      return;
    }
  }
}

int main() {
  std::mt19937 rng(12345);
  std::uniform_real_distribution<double> uni(0.0, 1.0);
  const int shapes[][3] = {{8, 8, 8}, {16, 16, 8}, {32, 16, 16}, {24, 24, 14}, {32, 32, 8},
                           {16, 16, 32}};
  std::vector<Case> cases;
  // circular orbits, magnification 1.1 .. 10
  const double mags[] = {1.1, 1.5, 2.0, 4.0, 10.0};
  for (double mag : mags) {
    const double R = 40.0 + 60.0 * uni(rng);
    cases.push_back(orbit(48, 44, 36, 56, 44, 9, R, R * (mag - 1.0), 0.0, 0.0, 0.0,
                          uni(rng) * 6.28));
  }
  // tilted Beijing-style detectors, source off the mid-plane
  for (int i = 0; i < 4; ++i) {
    const double R = 35.0 + 40.0 * uni(rng);
    cases.push_back(orbit(40, 40, 24, 48, 40, 8, R, R * (0.3 + 1.5 * uni(rng)),
                          0.35 * (uni(rng) - 0.5), 0.35 * (uni(rng) - 0.5),
                          10.0 * (uni(rng) - 0.5), uni(rng) * 6.28));
  }
  // sources next to the volume: half-diagonal 28, source at 30 .. 45
  for (int i = 0; i < 4; ++i) {
    const double R = 30.0 + 15.0 * uni(rng);
    cases.push_back(orbit(40, 40, 20, 64, 48, 8, R, 60.0, 0.0, 0.0, 4.0 * (uni(rng) - 0.5),
                          uni(rng) * 6.28));
  }
  for (const Case& cs : cases) {
    for (const auto& sh : shapes) {
      const int nbx = (cs.g.nx + sh[0] - 1) / sh[0], nby = (cs.g.ny + sh[1] - 1) / sh[1],
                nbz = (cs.g.nz + sh[2] - 1) / sh[2];
      // the corner bricks (edge bricks are partial), a middle one and two
      // seeded ones
      for (int q = 0; q < 5; ++q) {
        int jx, jy, jz;
        if (q < 2) {
          jx = q ? nbx - 1 : 0;
          jy = q ? nby - 1 : 0;
          jz = q ? nbz - 1 : 0;
        } else if (q == 2) {
          jx = nbx / 2;
          jy = nby / 2;
          jz = nbz / 2;
        } else {
          jx = (int)(uni(rng) * nbx);
          jy = (int)(uni(rng) * nby);
          jz = (int)(uni(rng) * nbz);
        }
        check_brick(cs, sh[0], sh[1], sh[2], jx, jy, jz, false);
        check_brick(cs, sh[0], sh[1], sh[2], jx, jy, jz, true);
        if (fails) return 1;
      }
    }
  }
  std::printf("ok %ld %ld %.6f\n", checks, inf_views, max_ratio);
  return 0;
}
