// Host check of dtv::exact_window (dtv-net_amd/csrc/dtv_geom.h): the run of
// samples of a 32.32 fixed-point ray whose floor index lies in a box.
//
// Reference run = brute force over every sample (geometry rays) or an exact
// per-axis binary search in int64 (synthetic far-away rays, where kend is too
// large to enumerate).  Prints "ok <cases> <hits>" or the first mismatch and
// exits 1.  Built and run by tests/test_window_native.py.
#include <cstdint>
#include <cstdio>
#include <random>

#include "dtv_geom.h"

using namespace dtv;

static bool member(const int64_t P[3], const int64_t D[3], const int lo[3], const int hi[3],
                   int64_t k) {
  for (int c = 0; c < 3; ++c) {
    const int64_t f = (P[c] + k * D[c]) >> 32;
    if (f < lo[c] || f > hi[c]) return false;
  }
  return true;
}

// exact bounds by binary search on each (monotone) axis
static void search_window(const int64_t P[3], const int64_t D[3], const int lo[3],
                          const int hi[3], int64_t kend, int64_t& ka, int64_t& kb) {
  int64_t a = 0, b = kend - 1;
  for (int c = 0; c < 3; ++c) {
    auto f = [&](int64_t k) { return (P[c] + k * D[c]) >> 32; };
    if (D[c] == 0) {
      if (f(0) < lo[c] || f(0) > hi[c]) b = a - 1;
      continue;
    }
    // first k in [0, kend) with the coordinate past the entry plane, last k
    // before the exit plane
    auto first = [&](auto pred) {
      int64_t l = 0, h = kend;  // pred monotone false->true
      while (l < h) {
        const int64_t m = l + (h - l) / 2;
        if (pred(m)) h = m; else l = m + 1;
      }
      return l;
    };
    int64_t s, e;
    if (D[c] > 0) {
      s = first([&](int64_t k) { return f(k) >= lo[c]; });
      e = first([&](int64_t k) { return f(k) > hi[c]; }) - 1;
    } else {
      s = first([&](int64_t k) { return f(k) <= hi[c]; });
      e = first([&](int64_t k) { return f(k) < lo[c]; }) - 1;
    }
    if (s > a) a = s;
    if (e < b) b = e;
  }
  if (b < a) { ka = kb = 0; } else { ka = a; kb = b + 1; }
}

static int fail(const char* what, long i, int64_t tka, int64_t tkb, int ka, int kb) {
  std::printf("mismatch (%s) case %ld: true [%lld, %lld) got [%d, %d)\n", what, i,
              (long long)tka, (long long)tkb, ka, kb);
  return 1;
}

int main() {
  std::mt19937_64 rng(12345);
  long cases = 0, hits = 0;
  // 1) rays of Standard (axis among x, y) and a circular table-like geometry,
  //    including the cardinal angles whose central column has a direction
  //    component far below 2^-33
  const double angles[] = {0.0, 90.0, 180.0, 270.0, 33.3, 137.9, 225.0, 301.5};
  for (int mode = 0; mode < 2; ++mode) {
    Geo g{};
    g.nx = 64; g.ny = 64; g.nz = 48; g.nu = 96; g.nv = 80;
    g.axis_mode = mode == 0 ? kAxisXY : kAxisXYZ;
    for (int c = 0; c < 3; ++c) {
      g.lo[c] = mode == 0 ? -0.5 * (&g.nx)[c] : -0.5;
      g.hi[c] = mode == 0 ? 1.5 * (&g.nx)[c] : (&g.nx)[c] - 0.5;
    }
    g.margin = mode == 0 ? 6.0 : 3.0;
    for (double ang : angles) {
      View v;
      view_standard(ang, 500.0, 1000.0, g.nx, g.ny, g.nz, g.nu, g.nv, v);
      for (int it = 0; it < 6000; ++it) {
        const int iu = (it % 7 == 0) ? g.nu / 2 : (int)(rng() % g.nu);
        const int iv = (int)(rng() % g.nv);
        double d[3], P0[3];
        int kend;
        ray_core(g, v, iu, iv, d, P0, kend);
        if (kend <= 0) continue;
        int64_t P[3], D[3];
        double Pd[3], Dd[3];
        for (int c = 0; c < 3; ++c) {
          P[c] = to_fix(P0[c]);
          D[c] = to_fix(d[c]);
          Pd[c] = rint(P0[c] * kFix);
          Dd[c] = rint(d[c] * kFix);
        }
        const int k0 = (int)(rng() % kend);
        int lo[3], hi[3];
        for (int c = 0; c < 3; ++c) {
          const int f = (int)((P[c] + (int64_t)k0 * D[c]) >> 32);
          lo[c] = f - (int)(rng() % 20);
          hi[c] = lo[c] + (int)(rng() % 40);
        }
        int64_t tka = -1, tkb = -1;
        for (int k = 0; k < kend; ++k)
          if (member(P, D, lo, hi, k)) {
            if (tka < 0) tka = k;
            tkb = k + 1;
          }
        for (int k = (int)tka; tka >= 0 && k < tkb; ++k)
          if (!member(P, D, lo, hi, k)) return fail("non-convex", cases, tka, tkb, 0, 0);
        int ka, kb;
        exact_window(Pd, Dd, lo, hi, kend, ka, kb);
        ++cases;
        if (tka < 0) {
          if (kb > ka) return fail("geometry", cases, 0, 0, ka, kb);
        } else {
          ++hits;
          if (ka != tka || kb != tkb) return fail("geometry", cases, tka, tkb, ka, kb);
        }
      }
    }
  }
  // 2) synthetic rays: start up to 2^19 voxels away, kend up to 2^20, some
  //    direction components exactly zero or below 2^-33, boxes on the ray
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  for (int it = 0; it < 200000; ++it) {
    double d[3];
    double n2 = 0.0;
    for (int c = 0; c < 3; ++c) {
      d[c] = U(rng);
      const int r = (int)(rng() % 10);
      if (r == 0) d[c] = 0.0;
      if (r == 1) d[c] = U(rng) * 1e-12;
      n2 += d[c] * d[c];
    }
    if (n2 < 1e-6) continue;
    for (int c = 0; c < 3; ++c) d[c] /= sqrt(n2);
    const int kend = 1 + (int)(rng() % (1u << 20));
    double P0[3];
    const double reach = (double)(rng() % (1u << 19));
    for (int c = 0; c < 3; ++c) P0[c] = U(rng) * reach;
    int64_t P[3], D[3];
    double Pd[3], Dd[3];
    for (int c = 0; c < 3; ++c) {
      P[c] = to_fix(P0[c]);
      D[c] = to_fix(d[c]);
      Pd[c] = rint(P0[c] * kFix);
      Dd[c] = rint(d[c] * kFix);
    }
    const int64_t k0 = (int64_t)(rng() % (uint64_t)kend);
    int lo[3], hi[3];
    for (int c = 0; c < 3; ++c) {
      const int f = (int)((P[c] + k0 * D[c]) >> 32);
      lo[c] = f - (int)(rng() % 24);
      hi[c] = lo[c] + (int)(rng() % 48);
    }
    int64_t tka, tkb;
    search_window(P, D, lo, hi, kend, tka, tkb);
    int ka, kb;
    exact_window(Pd, Dd, lo, hi, kend, ka, kb);
    ++cases;
    if (tkb <= tka) {
      if (kb > ka) return fail("synthetic", cases, tka, tkb, ka, kb);
    } else {
      ++hits;
      if (ka != tka || kb != tkb) return fail("synthetic", cases, tka, tkb, ka, kb);
      if (!member(P, D, lo, hi, ka) || !member(P, D, lo, hi, kb - 1) ||
          (ka > 0 && member(P, D, lo, hi, ka - 1)) || (kb < kend && member(P, D, lo, hi, kb)))
        return fail("synthetic-edges", cases, tka, tkb, ka, kb);
    }
  }
  std::printf("ok %ld %ld\n", cases, hits);
  return 0;
}
