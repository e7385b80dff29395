"""Brick kernels with a brick's views split over several workgroups
(``view_parts``, grid z): A's parts add their ray sums into the sinogram like
whole bricks, A^T's two parts add into a zero-filled volume.  Every split
meets the FP64 oracle at the projector bar, agrees with the unsplit launch to
FP32 rounding, and the two-part A^T is bitwise repeatable."""
import pytest
import torch

from conftest import rel_l2

from oracle import cone
from dtvproj import geometry as G, ops, phantom

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


CIRC = dict(vs=(30, 26, 18), vx=1.0, ds=(44, 30), pitch=1.2, sod=90.0, sdd=170.0, n=17)


def _circ():
    c = CIRC
    ang = G.uniform_angles(c["n"]) + 3.0
    g = G.circular(ang, c["vs"], c["vx"], c["ds"], c["pitch"], c["sod"], c["sdd"])
    args = (ang, c["vs"], c["vx"], c["ds"], c["pitch"], c["sod"], c["sdd"])
    return g, args


@pytest.mark.parametrize("brick", [(16, 16, 8), (26, 26, 14), (8, 8, 8)])
def test_forward_parts_meet_oracle(cuda, brick):
    """1 to 16 parts, including more parts than the 17 views split evenly
    (parts with one view or none)."""
    g, args = _circ()
    x = phantom.gaussian(g.volume_shape(2), 41)
    ref = cone.circ_forward(x, *args)
    one = ops.forward(x.to(cuda), g, dict(algorithm=2, brick=brick, block=256, view_parts=1))
    for p in (2, 3, 5, 16):
        _case("forward_parts", brick=brick, parts=p)
        out = ops.forward(x.to(cuda), g, dict(algorithm=2, brick=brick, block=256, view_parts=p))
        assert rel_l2(out, ref) <= TOL
        assert rel_l2(out, one) <= 1e-6


@pytest.mark.parametrize("partition", [False, True])
@pytest.mark.parametrize("brick", [(16, 16, 8), (32, 16, 16), (8, 8, 8)])
def test_adjoint_two_parts_meet_oracle_and_repeat(cuda, brick, partition):
    """Two parts adding into the zero-filled volume, through the
    unpartitioned and the partitioned bricks."""
    g, args = _circ()
    y = phantom.gaussian(g.sino_shape(2), 42)
    ref = cone.circ_adjoint(y, *args)
    _case("adjoint_parts", brick=brick, partition=partition)
    L2 = dict(algorithm=2, brick=brick, block=256, view_parts=2)
    one = ops.adjoint(y.to(cuda), g, dict(L2, view_parts=1), partition=partition)
    out = ops.adjoint(y.to(cuda), g, L2, partition=partition)
    again = ops.adjoint(y.to(cuda), g, L2, partition=partition)
    assert rel_l2(out, ref) <= TOL
    assert rel_l2(one, ref) <= TOL
    assert rel_l2(out, one) <= 2e-6
    assert torch.equal(out, again)


def test_culled_beijing_scan_with_parts(cuda):
    """147 narrow views: each brick culls the views it cannot reach before the
    parts divide what is left (the culled list, not the full scan)."""
    vs, ds = (40, 36, 16), (8, 24)
    P = G.beijing_default_vectors(n_angles=7, planes=21, sub_detector=ds, sid=70.0, sdd=114.0)
    g = G.beijing(P, vs, ds)
    vol = phantom.gaussian((1, 1, vs[2], vs[1], vs[0]), 43)
    ref = cone.bj_forward(vol, P, vs, ds)
    y = phantom.gaussian(g.sino_shape(1), 44)
    aref = cone.bj_adjoint(y, P, vs, ds)
    for p in (1, 4, 16):
        _case("beijing_culled_forward_parts", parts=p)
        out = ops.forward(vol.to(cuda), g, dict(algorithm=2, brick=(16, 16, 8), block=256,
                                                view_parts=p)).cpu()
        assert rel_l2(out, ref) <= TOL
        assert bool((out[ref == 0] == 0).all())
    for p in (1, 2):
        _case("beijing_culled_adjoint_parts", parts=p)
        adj = ops.adjoint(y.to(cuda), g, dict(algorithm=2, brick=(16, 16, 8), block=256,
                                              view_parts=p))
        assert rel_l2(adj, aref) <= TOL


def test_default_launch_splits_small_volumes(cuda):
    """The library's default launch (no launch struct) on a volume small
    enough to be split into many parts per brick: results meet the oracle and
    the default A^T is bitwise repeatable."""
    ang = G.uniform_angles(40)
    vs, ds = (48, 48, 40), (72, 56)
    g = G.circular(ang, vs, 1.0, ds, 1.0, 150.0, 300.0)
    args = (ang, vs, 1.0, ds, 1.0, 150.0, 300.0)
    _case("default_launch_small", vol=vs, views=40)
    x = phantom.gaussian(g.volume_shape(1), 45)
    assert rel_l2(ops.forward(x.to(cuda), g), cone.circ_forward(x, *args)) <= TOL
    y = phantom.gaussian(g.sino_shape(1), 46)
    a1 = ops.adjoint(y.to(cuda), g)
    a2 = ops.adjoint(y.to(cuda), g)
    assert rel_l2(a1, cone.circ_adjoint(y, *args)) <= TOL
    assert torch.equal(a1, a2)


def test_view_parts_out_of_range_rejected(cuda):
    """At most 16 parts; A^T beyond two parts goes through the workspace."""
    g, args = _circ()
    x = torch.zeros(g.volume_shape(1), device=cuda)
    y = phantom.gaussian(g.sino_shape(1), 51)
    with pytest.raises(RuntimeError):
        ops.forward(x, g, dict(algorithm=2, brick=(16, 16, 8), block=256, view_parts=17))
    with pytest.raises(RuntimeError):
        ops.adjoint(y.to(cuda), g, dict(algorithm=2, brick=(16, 16, 8), block=256, view_parts=17))
    three = ops.adjoint(y.to(cuda), g, dict(algorithm=2, brick=(16, 16, 8), block=256,
                                            view_parts=3))
    assert rel_l2(three, cone.circ_adjoint(y, *args)) <= TOL


@pytest.mark.parametrize("persistent", [0, 29])
def test_multi_instance_parts_match_single_calls(cuda, persistent):
    """Split bricks inside a multi-instance launch (one item per brick part)
    give what the single-instance calls give with the same split: A to FP32
    rounding, the two-part A^T bitwise (same views per part, integer groups,
    two commutative adds)."""
    from dtvproj import batched
    geoms = [G.circular(G.uniform_angles(13), (33, 29, 21), 1.0, (50, 36), 1.3, 120.0, 260.0),
             G.circular(G.uniform_angles(9) + 5.0, (24, 24, 30), 1.1, (40, 44), 1.4, 100.0, 230.0)]
    parts = [(4, 2), (7, 2)]
    bricks = [((16, 16, 8), (16, 16, 8)), ((24, 24, 14), (8, 8, 8))]
    _case("multi_parts_vs_single", persistent=persistent, parts=parts)
    mp = batched.MultiProjector(geoms, bricks=bricks, block=256, persistent_blocks=persistent,
                                view_parts=parts)
    vols = [phantom.gaussian(g.volume_shape(1), 47 + i).to(cuda) for i, g in enumerate(geoms)]
    ys = [phantom.gaussian(g.sino_shape(1), 49 + i).to(cuda) for i, g in enumerate(geoms)]
    sinos = mp.forward(vols)
    outs = mp.adjoint(ys)
    for i, g in enumerate(geoms):
        f = ops.forward(vols[i], g, dict(algorithm=2, brick=bricks[i][0], block=256,
                                         view_parts=parts[i][0]))
        assert rel_l2(sinos[i], f) <= 1e-6
        a = ops.adjoint(ys[i], g, dict(algorithm=2, brick=bricks[i][1], block=256,
                                       view_parts=parts[i][1]))
        assert torch.equal(outs[i], a)


def test_auto_parts_split_the_costly_bricks():
    """The multi launch's automatic split: instances whose bricks cost more
    than the launch's target item get more parts, A^T at most two."""
    from dtvproj import batched, configs
    geoms = configs.uneven_16()
    bricks = [((24, 24, 14), (32, 16, 16))] * len(geoms)
    parts = batched.auto_parts(geoms, bricks, 768)
    assert len(parts) == len(geoms)
    assert all(1 <= f <= 16 and a in (1, 2) for f, a in parts)
    cost = [g.n_views for g in geoms]
    hi, lo = max(range(len(geoms)), key=cost.__getitem__), min(range(len(geoms)), key=cost.__getitem__)
    assert parts[hi][0] >= parts[lo][0]


def _adjoint_ws(y, g, cuda, brick, parts):
    """A^T through dtv_cone_adjoint_ws with an explicit split (workspace size
    from dtv_cone_adjoint_plan)."""
    import ctypes
    from dtvproj import _capi
    B = y.shape[0]
    out = torch.empty(g.volume_shape(B), device=cuda)
    L = _capi.launch_struct(algorithm=2, brick=brick, block=256, view_parts=parts)
    _, nb = ops.adjoint_plan(g, B, dict(algorithm=2, brick=brick, block=256, view_parts=parts))
    ws = torch.empty(max(nb // 4, 1), device=cuda)
    sc = _capi.scheme_struct(g.scheme)
    rc = _capi.lib().dtv_cone_adjoint_ws(
        y.data_ptr(), out.data_ptr(), B, _capi.int3(g.volume_size), _capi.int2(g.detector_size),
        g.device_views(cuda).data_ptr(), g.n_views, ctypes.byref(sc), ctypes.byref(L),
        ws.data_ptr(), nb, ops._stream(cuda))
    _capi.check(rc, "dtv_cone_adjoint_ws")
    return out


@pytest.mark.parametrize("brick", [(16, 16, 8), (32, 16, 16), (8, 8, 8)])
def test_adjoint_workspace_split_meets_oracle_and_repeats(cuda, brick):
    """Parts that store partial volumes in a workspace, added up in part
    order by a second kernel: 1 to 16 parts (more than the 17 views split
    evenly) meet the oracle, agree with the unsplit launch and repeat
    bitwise."""
    g, args = _circ()
    y = phantom.gaussian(g.sino_shape(2), 48)
    ref = cone.circ_adjoint(y, *args)
    yd = y.to(cuda)
    one = _adjoint_ws(yd, g, cuda, brick, 1)
    assert rel_l2(one, ref) <= TOL
    for p in (2, 3, 7, 16):
        _case("adjoint_ws_parts", brick=brick, parts=p)
        out = _adjoint_ws(yd, g, cuda, brick, p)
        assert rel_l2(out, ref) <= TOL
        assert rel_l2(out, one) <= 2e-6
        assert torch.equal(out, _adjoint_ws(yd, g, cuda, brick, p))


def test_adjoint_workspace_split_culled_beijing(cuda):
    vs, ds = (40, 36, 16), (8, 24)
    P = G.beijing_default_vectors(n_angles=7, planes=21, sub_detector=ds, sid=70.0, sdd=114.0)
    g = G.beijing(P, vs, ds)
    y = phantom.gaussian(g.sino_shape(1), 49)
    aref = cone.bj_adjoint(y, P, vs, ds)
    for p in (4, 16):
        _case("adjoint_ws_beijing_culled", parts=p)
        assert rel_l2(_adjoint_ws(y.to(cuda), g, cuda, (16, 16, 8), p), aref) <= TOL


def test_default_adjoint_uses_the_planned_split(cuda):
    """ops.adjoint without a launch runs the library's plan (here many parts
    through a workspace) and matches an explicit call of that plan bitwise."""
    ang = G.uniform_angles(40)
    vs, ds = (48, 48, 40), (72, 56)
    g = G.circular(ang, vs, 1.0, ds, 1.0, 150.0, 300.0)
    plan, nbytes = ops.adjoint_plan(g, 1)
    assert plan.view_parts > 2 and nbytes > 0
    y = phantom.gaussian(g.sino_shape(1), 50).to(cuda)
    _case("default_adjoint_plan", parts=plan.view_parts, brick=tuple(plan.brick))
    assert torch.equal(ops.adjoint(y, g),
                       _adjoint_ws(y, g, cuda, tuple(plan.brick), plan.view_parts))


@pytest.mark.parametrize("launch", [None, dict(algorithm=2, brick=(16, 16, 8), block=256,
                                               view_parts=5),
                                    dict(algorithm=2, brick=(26, 26, 14), block=256)])
def test_deterministic_forward_repeats_bitwise_and_meets_oracle(cuda, launch):
    """The int64 fixed-point forward (dtv_cone_forward_ws): bitwise equal on
    repeated calls, within FP32 rounding of the atomic version, and at the
    oracle bar; a batch of two."""
    g, args = _circ()
    x = phantom.gaussian(g.volume_shape(2), 52)
    xd = x.to(cuda)
    _case("deterministic_forward", launch=launch)
    a = ops.forward(xd, g, launch, deterministic=True)
    b = ops.forward(xd, g, launch, deterministic=True)
    assert torch.equal(a, b)
    assert rel_l2(a, cone.circ_forward(x, *args)) <= TOL
    assert rel_l2(a, ops.forward(xd, g, launch, deterministic=False)) <= 1e-6


def test_deterministic_forward_follows_torch_flag_and_keeps_nonfinite(cuda):
    """torch.use_deterministic_algorithms(True) selects the fixed-point path;
    a volume holding inf or NaN takes the FP32 path, so they reach exactly the
    rays they reach on it."""
    g, args = _circ()
    x = phantom.gaussian(g.volume_shape(1), 53).to(cuda)
    prev = torch.are_deterministic_algorithms_enabled()
    try:
        torch.use_deterministic_algorithms(True)
        a, b = ops.forward(x, g), ops.forward(x, g)
    finally:
        torch.use_deterministic_algorithms(prev)
    assert torch.equal(a, b)
    assert torch.equal(a, ops.forward(x, g, deterministic=True))
    xi = x.clone()
    xi[0, 0, 9, 13, 15] = float("inf")
    xi[0, 0, 3, 4, 5] = float("nan")
    fi = ops.forward(xi, g, deterministic=True)
    ref = ops.forward(xi, g, dict(algorithm=1))   # ray-driven gather
    assert torch.equal(torch.isfinite(fi), torch.isfinite(ref))
    assert torch.equal(torch.isnan(fi), torch.isnan(ref))
