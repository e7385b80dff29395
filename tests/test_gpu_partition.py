"""GPU: the partitioned matched A^T (``dtv_cone_adjoint_ws``, the default of
``ops.adjoint``).  Each sample is marched by the one brick holding its floor
cell; taps on a brick's upper halo are summed per voxel and added to the
neighbour's voxels by a second kernel in a fixed order.  Checked against the
FP64 oracle on volumes whose sizes are not multiples of the brick (partial
bricks on every upper face, floor -1 samples on every lower face), on brick
shapes with and without compile-time kernels, over 1 to 16 view parts and a
batch of two; against the unpartitioned kernel to FP32 rounding; bitwise
repeatable; non-finite values reach the same voxels as in the unpartitioned
kernel."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

from oracle import cone
from dtvproj import geometry as G, ops, phantom

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _ws_launch(brick, parts):
    return dict(algorithm=2, brick=brick, block=256, view_parts=parts)


def _adjoint_ws(y, g, brick, parts):
    """A^T through the workspace entry point with a fixed brick and split."""
    import ctypes
    from dtvproj import _capi
    plan, nb = ops.adjoint_plan(g, y.shape[0], _ws_launch(brick, parts))
    assert nb > 0
    out = torch.empty(g.volume_shape(y.shape[0]), dtype=torch.float32, device=y.device)
    ws = torch.empty(nb // 4, dtype=torch.float32, device=y.device)
    sc = _capi.scheme_struct(g.scheme)
    rc = _capi.lib().dtv_cone_adjoint_ws(
        y.data_ptr(), out.data_ptr(), y.shape[0], _capi.int3(g.volume_size),
        _capi.int2(g.detector_size), g.device_views(y.device).data_ptr(), g.n_views,
        ctypes.byref(sc), ctypes.byref(plan), ws.data_ptr(), nb, ops._stream(y.device))
    _capi.check(rc, "dtv_cone_adjoint_ws")
    return out


# 45 x 38 x 29: no axis a multiple of any brick edge below
CIRC = dict(vs=(45, 38, 29), vx=0.9, ds=(64, 40), pitch=1.3, sod=80.0, sdd=160.0, n=13)


def _circ():
    c = CIRC
    ang = G.uniform_angles(c["n"]) + 7.0
    g = G.circular(ang, c["vs"], c["vx"], c["ds"], c["pitch"], c["sod"], c["sdd"])
    return g, (ang, c["vs"], c["vx"], c["ds"], c["pitch"], c["sod"], c["sdd"])


SHAPES = [(32, 16, 16), (16, 16, 8), (24, 24, 14), (8, 8, 8), (7, 5, 3), (12, 10, 6)]


@pytest.mark.parametrize("brick", SHAPES, ids=lambda b: "x".join(map(str, b)))
def test_partitioned_adjoint_meets_oracle(cuda, brick):
    g, args = _circ()
    y = phantom.gaussian(g.sino_shape(2), 61)
    ref = cone.circ_adjoint(y, *args)
    yd = y.to(cuda)
    classic = ops.adjoint(yd, g, _ws_launch(brick, 1), partition=False)
    assert rel_l2(classic, ref) <= TOL
    for p in (1, 2, 5, 16):
        _case("partitioned_adjoint", brick=brick, parts=p)
        out = _adjoint_ws(yd, g, brick, p)
        assert rel_l2(out, ref) <= TOL
        assert rel_l2(out, classic) <= 2e-6
        assert torch.equal(out, _adjoint_ws(yd, g, brick, p))


def test_partitioned_adjoint_single_ray_and_faces(cuda):
    """One non-zero detector pixel: the adjoint is that ray's trilinear
    weights, all of them on voxels the ray passes, wherever its samples fall
    relative to the bricks' faces; the partitioned and unpartitioned kernels
    agree on every voxel to FP32 rounding and the oracle's support."""
    g, args = _circ()
    for (iu, iv, a) in ((31, 19, 0), (0, 0, 4), (63, 39, 9), (40, 3, 12)):
        y = torch.zeros(g.sino_shape(1), dtype=torch.float64)
        y[0, 0, a, iv, iu] = 1.0
        ref = cone.circ_adjoint(y, *args)
        _case("single_ray", pixel=(iu, iv), view=a)
        out = _adjoint_ws(y.float().to(cuda), g, (16, 16, 8), 1).cpu().double()
        assert rel_l2(out, ref) <= TOL
        assert bool((out[ref == 0] == 0).all())


def test_partitioned_adjoint_standard_and_beijing(cuda):
    """The Standard scheme (samples phased from the world-unit box) and a
    147-view Beijing scan whose bricks cull most views, through the default
    (partitioned) ops.adjoint, against the oracle and the unpartitioned
    kernel."""
    vs, ds = (48, 40, 24), (60, 20)
    ang = G.uniform_angles(16)
    g = G.standard(ang, vs, ds, 100.0, 200.0)
    y = phantom.gaussian(g.sino_shape(1), 62)
    ref = cone.std_adjoint(y, ang, vs, ds, 100.0, 200.0)
    _case("standard_default")
    out = ops.adjoint(y.to(cuda), g)
    assert rel_l2(out, ref) <= TOL
    assert rel_l2(out, ops.adjoint(y.to(cuda), g, _ws_launch((32, 16, 16), 1),
                                   partition=False)) <= 2e-6
    vs, ds = (40, 36, 16), (8, 24)
    P = G.beijing_default_vectors(n_angles=7, planes=21, sub_detector=ds, sid=70.0, sdd=114.0)
    gb = G.beijing(P, vs, ds)
    yb = phantom.gaussian(gb.sino_shape(1), 63)
    refb = cone.bj_adjoint(yb, P, vs, ds)
    for brick, p in (((16, 16, 8), 1), ((16, 16, 8), 4), ((32, 16, 16), 3)):
        _case("beijing_partitioned", brick=brick, parts=p)
        assert rel_l2(_adjoint_ws(yb.to(cuda), gb, brick, p), refb) <= TOL


def test_partitioned_adjoint_near_source(cuda):
    """A source close to the volume (large depth ratios across a brick: some
    views take the FP32 groups) on odd-sized bricks."""
    vs, ds = (37, 29, 21), (45, 27)
    ang = G.uniform_angles(9) + 2.0
    g = G.circular(ang, vs, 0.9, ds, 1.1, 60.0, 130.0)
    y = phantom.gaussian(g.sino_shape(2), 64)
    ref = cone.circ_adjoint(y, ang, vs, 0.9, ds, 1.1, 60.0, 130.0)
    for brick in ((16, 16, 8), (32, 16, 16), (7, 5, 3)):
        _case("near_source_partitioned", brick=brick)
        out = _adjoint_ws(y.to(cuda), g, brick, 3)
        assert rel_l2(out, ref) <= TOL
        assert rel_l2(out, ops.adjoint(y.to(cuda), g, _ws_launch(brick, 1),
                                       partition=False)) <= 5e-6


def test_partitioned_adjoint_non_finite_values(cuda):
    """inf and NaN in the sinogram reach the same voxels as through the
    unpartitioned kernel (NaN where a tap of such a ray lands, the finite
    voxels equal to FP32 rounding)."""
    g, _ = _circ()
    y = phantom.gaussian(g.sino_shape(1), 65)
    y[0, 0, 2, 20, 30] = float("inf")
    y[0, 0, 7, 5, 12] = float("nan")
    y[0, 0, 11, 39, 0] = -float("inf")
    yd = y.to(cuda)
    for brick in ((16, 16, 8), (32, 16, 16)):
        _case("non_finite_partitioned", brick=brick)
        classic = ops.adjoint(yd, g, _ws_launch(brick, 1), partition=False).cpu()
        out = _adjoint_ws(yd, g, brick, 2).cpu()
        bad = ~torch.isfinite(classic)
        assert int(bad.sum()) > 0
        assert torch.equal(bad, ~torch.isfinite(out))
        fin = ~bad
        assert rel_l2(out[fin], classic[fin]) <= 2e-6


def test_default_adjoint_is_partitioned_and_repeatable(cuda):
    """ops.adjoint with no launch plans a workspace (halo sums included) and
    runs the partitioned kernel: equal to the explicit workspace call with
    the planned shape and split, and bitwise repeatable."""
    g, args = _circ()
    plan, nbytes = ops.adjoint_plan(g, 1)
    bx, by, bz = plan.brick
    X, Y, Z = g.volume_size
    nb = -(-X // bx) * -(-Y // by) * -(-Z // bz)
    halo = (bx + 1) * (by + 1) * (bz + 1) - bx * by * bz
    assert nbytes == ((plan.view_parts - 1) * X * Y * Z + plan.view_parts * nb * halo) * 4
    y = phantom.gaussian(g.sino_shape(1), 66).to(cuda)
    _case("default_partitioned", brick=tuple(plan.brick), parts=plan.view_parts)
    out = ops.adjoint(y, g)
    assert torch.equal(out, _adjoint_ws(y, g, tuple(plan.brick), plan.view_parts))
    assert torch.equal(out, ops.adjoint(y, g))
    assert rel_l2(out, cone.circ_adjoint(y.cpu().double(), *args)) <= TOL
