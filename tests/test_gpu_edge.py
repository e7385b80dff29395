"""GPU edge cases of the projector pair: non-finite values, zero and empty
inputs, rays that miss the volume, degenerate volume shapes, one view,
batches, and a source inside the volume's bounding box (brick corners behind
the source, which sends the brick kernels down their whole-detector path).

Every case runs both kernel families (ray-driven and brick) and, where the
sizes allow, the FP64 oracle."""
import math

import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import cone
from dtvproj import geometry as G, ops, phantom

pytestmark = pytest.mark.gpu

RAY = dict(algorithm=1)
BRICK = dict(algorithm=2, brick=(8, 8, 8), block=64)
DEFAULT = None
LAUNCHES = [DEFAULT, RAY, BRICK]
TOL = 1e-5


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _circ(vs, n_views=9, det=(24, 20), vx=1.0, pitch=1.5, sod=80.0, sdd=160.0, offset=0.0):
    ang = G.uniform_angles(n_views) + offset
    return G.circular(ang, vs, vx, det, pitch, sod, sdd), ang, (vs, vx, det, pitch, sod, sdd)


@pytest.mark.parametrize("launch", LAUNCHES)
def test_nonfinite_values_propagate(cuda, launch):
    _case("nonfinite", launch=launch)
    g, ang, args = _circ((12, 10, 8))
    y = phantom.gaussian(g.sino_shape(1), 41)
    V, U = g.detector_size[1], g.detector_size[0]
    y[0, 0, 2, V // 2, U // 2] = float("inf")
    y[0, 0, 5, V // 2 + 1, U // 2 - 2] = float("nan")
    a = ops.adjoint(y.to(cuda), g, launch).cpu()
    ref = ops.adjoint(y.to(cuda), g, RAY).cpu()
    # the same voxels are non-finite, the rest agree
    assert torch.equal(torch.isfinite(a), torch.isfinite(ref))
    assert (~torch.isfinite(a)).any()
    m = torch.isfinite(ref)
    assert rel_l2(a[m], ref[m]) <= 5e-6
    x = phantom.gaussian(g.volume_shape(1), 42)
    x[0, 0, 4, 5, 6] = float("nan")
    f = ops.forward(x.to(cuda), g, launch).cpu()
    fr = ops.forward(x.to(cuda), g, RAY).cpu()
    assert torch.equal(torch.isnan(f), torch.isnan(fr)) and torch.isnan(f).any()


@pytest.mark.parametrize("launch", LAUNCHES)
def test_zero_inputs_give_exact_zeros(cuda, launch):
    _case("zeros", launch=launch)
    g, _, _ = _circ((12, 10, 8))
    assert torch.count_nonzero(ops.forward(torch.zeros(g.volume_shape(2), device=cuda), g, launch)) == 0
    assert torch.count_nonzero(ops.adjoint(torch.zeros(g.sino_shape(2), device=cuda), g, launch)) == 0


@pytest.mark.parametrize("launch", LAUNCHES)
def test_detector_missing_the_volume(cuda, launch):
    """A detector shifted far off-axis: no ray meets the volume."""
    _case("missing detector", launch=launch)
    g, _, _ = _circ((12, 10, 8))
    views = g.views.copy()
    views[:, 3:6] += 400.0 * views[:, 9:12] / np.linalg.norm(views[:, 9:12], axis=1, keepdims=True)
    g2 = G.ConeGeometry(g.volume_size, g.detector_size, views, g.scheme)
    x = phantom.gaussian(g2.volume_shape(1), 43).to(cuda)
    y = phantom.gaussian(g2.sino_shape(1), 44).to(cuda)
    assert torch.count_nonzero(ops.forward(x, g2, launch)) == 0
    assert torch.count_nonzero(ops.adjoint(y, g2, launch)) == 0


@pytest.mark.parametrize("vs", [(1, 1, 1), (1, 7, 1), (5, 1, 3), (2, 3, 33)])
@pytest.mark.parametrize("launch", LAUNCHES)
def test_degenerate_volume_shapes(cuda, vs, launch):
    _case("degenerate volume", vs=vs, launch=launch)
    g, ang, args = _circ(vs, n_views=7, det=(16, 40), pitch=1.0, sod=60.0, sdd=120.0)
    X, Y, Z = vs
    x = phantom.gaussian((1, 1, Z, Y, X), 45)
    y = phantom.gaussian(g.sino_shape(1), 46)
    assert rel_l2(ops.forward(x.to(cuda), g, launch), cone.circ_forward(x, ang, *args)) <= TOL
    assert rel_l2(ops.adjoint(y.to(cuda), g, launch), cone.circ_adjoint(y, ang, *args)) <= TOL


@pytest.mark.parametrize("launch", LAUNCHES)
def test_single_view_and_batch(cuda, launch):
    _case("single view / batch", launch=launch)
    g, ang, args = _circ((14, 12, 9), n_views=1, offset=33.0)
    x = phantom.gaussian((3,) + g.volume_shape(1)[1:], 47)
    y = phantom.gaussian((3,) + g.sino_shape(1)[1:], 48)
    f = ops.forward(x.to(cuda), g, launch)
    a = ops.adjoint(y.to(cuda), g, launch)
    assert rel_l2(f, cone.circ_forward(x, ang, *args)) <= TOL
    assert rel_l2(a, cone.circ_adjoint(y, ang, *args)) <= TOL
    # batch items are independent: each equals its own single call
    for b in range(3):
        assert rel_l2(f[b:b + 1], ops.forward(x[b:b + 1].to(cuda), g, launch)) <= 1e-6
        assert rel_l2(a[b:b + 1], ops.adjoint(y[b:b + 1].to(cuda), g, launch)) <= 1e-6


@pytest.mark.parametrize("launch", LAUNCHES)
def test_source_inside_the_bounding_box(cuda, launch):
    """SOD below the volume's half-diagonal: some brick corners lie behind
    the source for some views (whole-detector candidate rectangles)."""
    _case("source inside bbox", launch=launch)
    vs = (24, 24, 10)
    half_diag = math.hypot(12, 12)
    assert 12.0 < 14.0 < half_diag
    g, ang, args = _circ(vs, n_views=6, det=(40, 30), pitch=2.0, sod=14.0,
                         sdd=60.0, offset=10.0)
    x = phantom.gaussian((1, 1, 10, 24, 24), 49)
    y = phantom.gaussian(g.sino_shape(1), 50)
    assert rel_l2(ops.forward(x.to(cuda), g, launch), cone.circ_forward(x, ang, *args)) <= TOL
    assert rel_l2(ops.adjoint(y.to(cuda), g, launch), cone.circ_adjoint(y, ang, *args)) <= TOL


@pytest.mark.parametrize("sid", [8.3, 40.0])
def test_beijing_backprojector_skipped_views(cuda, sid):
    """The Beijing reference backprojector skips, per workgroup, the views
    whose narrow sub-detector its voxels cannot reach; with 21 planes of 6
    pixels most views are skipped for most workgroups, and with the source
    inside the volume (sid 8.3, off the voxel lattice) the denominator
    changes sign inside workgroups.  Either way the result is the oracle's."""
    _case("beijing bp_ref skipped views", sid=sid)
    vs, ds = (70, 20, 18), (6, 9)
    P = G.beijing_default_vectors(3, 21, ds, sid, sid + 40.0, (0.7, 0.8))
    g = G.beijing(P, vs, ds)
    y = phantom.gaussian(g.sino_shape(1), 51)
    got = ops.bp_ref(y.to(cuda), g)
    want = cone.bj_bp_ref(y, P, vs, ds)
    assert rel_l2(got, want) <= TOL


def test_outputs_accumulated_with_atomics_must_be_device_memory(cuda):
    """The brick A and the ray-driven A^T add into their outputs with
    hardware FP32 atomics, which are not performed on host or managed
    memory: such an output is rejected with DTV_ERR_INVALID_ARGUMENT
    instead of giving silently wrong sums."""
    import ctypes
    from dtvproj import _capi, configs, ops
    g = configs.fd_32()
    print("[case] atomic outputs must be device memory", flush=True)
    L = _capi.lib()
    vol = torch.zeros(g.volume_shape(1), device=cuda)
    sino = torch.zeros(g.sino_shape(1), device=cuda)
    views = g.device_views(cuda)
    sc = _capi.scheme_struct(g.scheme)
    host_sino = torch.zeros(g.sino_shape(1)).pin_memory()
    host_vol = torch.zeros(g.volume_shape(1)).pin_memory()
    brick = _capi.launch_struct(2, (16, 16, 8), 256)
    ray = _capi.launch_struct(1)
    args = (1, _capi.int3(g.volume_size), _capi.int2(g.detector_size), views.data_ptr(),
            g.n_views, ctypes.byref(sc))
    st = ops._stream(cuda)
    assert L.dtv_cone_forward_ex(vol.data_ptr(), sino.data_ptr(), *args, ctypes.byref(brick),
                                 st) == 0
    assert L.dtv_cone_forward_ex(vol.data_ptr(), host_sino.data_ptr(), *args,
                                 ctypes.byref(brick), st) == 1
    assert L.dtv_cone_adjoint_ex(sino.data_ptr(), vol.data_ptr(), *args, ctypes.byref(ray),
                                 st) == 0
    assert L.dtv_cone_adjoint_ex(sino.data_ptr(), host_vol.data_ptr(), *args,
                                 ctypes.byref(ray), st) == 1
    torch.cuda.synchronize()
    # the rejection leaves no error behind for the next call
    assert torch.isfinite(ops.forward(vol, g)).all()
