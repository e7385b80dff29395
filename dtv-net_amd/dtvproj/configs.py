"""The benchmark workloads as geometries (physical units, mm).

Config names follow what they are:

* ``fista_64``     -- 64^3 at 1 mm, 72 views over 360 deg, 128x128 detector
  at 1.0 mm, SOD 500, SDD 1000 (plain pair + 10 FISTA-TV iterations).
* ``sparse_256``   -- 256^3 at 0.5 mm, 60 views, 384x384 at 0.8 mm, SOD 600,
  SDD 1100 (headline A + A^T throughput).
* ``unroll_128``   -- batch 4 of 128^3 at 1 mm, 90 views, 192x192 at 1.0 mm;
  the source distances are not stated for it, so it reuses SOD 500 / SDD
  1000 (magnification 2: the 96 mm detector field truncates the 128 mm
  volume laterally, as the configuration implies).
* ``fd_32``        -- 32^3 / 24 views / 48x48 finite-difference instance.
* ``large_512``    -- 512x512x384 at 0.4 mm, 120 views, 768x576 at 0.6 mm,
  SOD 700, SDD 1200.
* ``uneven_16``    -- 16 instances, X=Y in [64, 256], Z in [64, 160],
  views in [30, 120] (seeded), detectors scaled to cover the field of view.
"""
from __future__ import annotations

import math
import random

from . import geometry as G


def fista_64():
    return G.circular(G.uniform_angles(72), (64, 64, 64), 1.0, (128, 128), 1.0, 500.0, 1000.0)


def sparse_256():
    return G.circular(G.uniform_angles(60), (256, 256, 256), 0.5, (384, 384), 0.8, 600.0, 1100.0)


def unroll_128():
    return G.circular(G.uniform_angles(90), (128, 128, 128), 1.0, (192, 192), 1.0, 500.0, 1000.0)


def fd_32():
    return G.circular(G.uniform_angles(24), (32, 32, 32), 1.0, (48, 48), 1.0, 500.0, 1000.0)


def large_512():
    return G.circular(G.uniform_angles(120), (512, 512, 384), 0.4, (768, 576), 0.6, 700.0, 1200.0)


def uneven_16(seed: int = 4, n: int = 16):
    """Seeded list of ``n`` geometries of uneven size and view count."""
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        xy = rng.randint(64, 256)
        z = rng.randint(64, 160)
        views = rng.randint(30, 120)
        vx = 1.0
        sod, sdd = 500.0, 1000.0
        mag = sdd / sod
        # detector wide enough for the rotated field (diagonal) and the height
        # seen from the source, at pitch 1 mm x magnification
        fov_u = math.sqrt(2.0) * xy * vx * mag * 1.05
        fov_v = (z * vx + xy * vx / math.sqrt(2.0) * z / (sod)) * mag * 1.05
        pitch = 1.0 * mag / 1.0
        nu = int(math.ceil(fov_u / pitch))
        nv = int(math.ceil(fov_v / pitch))
        out.append(G.circular(G.uniform_angles(views), (xy, xy, z), vx, (nu, nv), pitch, sod, sdd))
    return out
