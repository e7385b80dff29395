"""CPU: the brick kernels' pass-1 line test (FP32, the brick's sample box
widened by 1/16 voxel) never drops a ray that has a sample in the box.

The FP32 test is restated here in numpy float32 as the kernels evaluate it
(view vectors rounded to FP32, slab test on the line through pixel iu, iv);
which rays really have a sample whose floor index lies in the box comes from
the oracle's FP64 ray enumeration.  Geometries: the 256^3 / 60-view
workload, the Beijing layout, a source 40 voxels from the centre of a 64^3
volume (inside its circumscribed sphere) and a source 10^4 voxels away."""
import numpy as np
import pytest

from oracle import cone
from dtvproj import configs, geometry as G

SLACK = np.float32(0.0625)  # the kernels' kLineSlack


def _line_hits(views, iu, iv, blo, bhi, rcp_ulp=False):
    """The kernels' FP32 slab test.  The kernels take 1/rc from v_rcp_f32,
    whose result is within 1 ulp of the correctly rounded reciprocal (AMD
    CDNA ISA, V_RCP_F32); with ``rcp_ulp`` every axis's reciprocal may be
    the neighbour on either side, and a ray counts as a hit only if it hits
    for every such choice (the adversarial combination: the largest entry
    and the smallest exit over both neighbours)."""
    fv = views.astype(np.float32)
    tin = np.full(iu.shape, -3.0e38, np.float32)
    tout = np.full(iu.shape, 3.0e38, np.float32)
    for c in range(3):
        rc = (fv[:, 3 + c] + iu * fv[:, 6 + c]) + iv * fv[:, 9 + c] - fv[:, c]
        rc = np.where(rc != 0, rc, np.float32(1e-30)).astype(np.float32)
        ir = (np.float32(1.0) / rc).astype(np.float32)
        irs = [ir]
        if rcp_ulp:
            irs = [np.nextafter(ir, np.float32(-np.inf)), ir,
                   np.nextafter(ir, np.float32(np.inf))]
        for r in irs:
            t0 = ((blo[c] - fv[:, c]) * r).astype(np.float32)
            t1 = ((bhi[c] - fv[:, c]) * r).astype(np.float32)
            tin = np.maximum(tin, np.minimum(t0, t1))
            tout = np.minimum(tout, np.maximum(t0, t1))
    return tout >= tin


def _rays_with_samples(g, sel, flo, fhi):
    """FP64: which of the selected rays have a sample whose floor index lies in
    [flo, fhi] on every axis."""
    sc = g.scheme
    r = cone.table_rays(g.views, g.detector_size, sc.axis_mode, sc.box_lo, sc.box_hi, sc.margin,
                        sc.scale)
    src = r.src.numpy()[sel]
    d = r.d.numpy()[sel]
    a0 = r.a_start.reshape(-1).numpy()[sel]
    a1 = r.a_end.reshape(-1).numpy()[sel]
    kmax = int(np.ceil((a1 - a0).max())) + 1
    k = np.arange(kmax, dtype=np.float64)
    alpha = a0[:, None] + k[None, :]
    valid = alpha < a1[:, None]
    inside = valid
    for c in range(3):
        f = np.floor(src[:, c, None] + alpha * d[:, c, None])
        inside = inside & (f >= flo[c]) & (f <= fhi[c])
    return inside.any(axis=1)


def _geometries():
    ang = G.uniform_angles(12)
    return {
        "sparse_256": configs.sparse_256().subset([0, 10, 20, 30, 40, 50]),
        "beijing": G.beijing(G.beijing_default_vectors(3, 21, (20, 24), 70.0, 114.0,
                                                       (1.0, 1.0)), (40, 36, 16), (20, 24)),
        "near_source": G.circular(ang, (64, 64, 64), 1.0, (96, 96), 1.0, 40.0, 120.0),
        "far_source": G.circular(ang, (64, 64, 64), 0.1, (96, 96), 0.25, 1000.0, 2000.0),
    }


@pytest.mark.parametrize("name", ["sparse_256", "beijing", "near_source", "far_source"])
def test_pass1_line_test_keeps_every_ray_with_a_sample(name):
    g = _geometries()[name]
    rng = np.random.default_rng(7)
    U, V = g.detector_size
    nv = g.n_views
    # a random subset of rays (all of them on the small detectors)
    n = nv * U * V
    sel = np.sort(rng.choice(n, size=min(n, 30000), replace=False))
    a, rem = np.divmod(sel, U * V)
    iv, iu = np.divmod(rem, U)
    views = np.asarray(g.views)[a]
    iu = iu.astype(np.float32)
    iv = iv.astype(np.float32)
    N = g.volume_size
    missed_at_negative = 0
    for trial in range(8):
        b = [(24, 24, 14), (32, 16, 16), (16, 16, 8), (26, 26, 14)][trial % 4]
        # the forward's box: floor(p) in [o, o+b-1], o in [-1, N-1]
        o = [int(rng.integers(-1, max(0, N[c] - b[c] // 2))) for c in range(3)]
        flo = o
        fhi = [o[c] + min(b[c], N[c] - o[c]) - 1 for c in range(3)]
        has = _rays_with_samples(g, sel, flo, fhi)
        if not has.any():
            continue
        blo = np.array([flo[c] - SLACK for c in range(3)], np.float32)
        bhi = np.array([fhi[c] + 1 + SLACK for c in range(3)], np.float32)
        hit = _line_hits(views, iu, iv, blo, bhi)
        assert not (has & ~hit).any(), (name, trial, int((has & ~hit).sum()))
        # and with v_rcp_f32's 1-ulp reciprocal, adversarially on every axis
        hit_rcp = _line_hits(views, iu, iv, blo, bhi, rcp_ulp=True)
        assert not (has & ~hit_rcp).any(), (name, trial, int((has & ~hit_rcp).sum()))
        # the check has teeth: a box shrunk by half a voxel drops such rays
        hit_in = _line_hits(views, iu, iv, blo + np.float32(0.5625), bhi - np.float32(0.5625))
        missed_at_negative += int((has & ~hit_in).sum())
    assert missed_at_negative > 0
