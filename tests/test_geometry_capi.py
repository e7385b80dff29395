"""CPU: product geometry vs the oracle, the C ABI library, host-side error
paths and the algorithmic byte accounting (no GPU compute)."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

from oracle import cone
from dtvproj import _capi, configs, geometry as G, ops, roofline

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _table_forward(g, vol):
    sc = g.scheme
    return cone.table_forward(vol, g.views, g.volume_size, g.detector_size, sc.axis_mode,
                              sc.box_lo, sc.box_hi, sc.margin, sc.scale)


def test_standard_table_matches_native_units():
    vs, ds = (9, 8, 7), (11, 6)
    ang = np.linspace(0, 360, 8, False) + 1.5
    vol = torch.from_numpy(np.random.default_rng(1).standard_normal((1, 1, 7, 8, 9)))
    g = G.standard(ang, vs, ds, 10.0, 20.0)
    a = cone.std_forward(vol, ang, vs, ds, 10.0, 20.0)
    assert float((a - _table_forward(g, vol)).abs().max()) < 1e-12


def test_beijing_table_matches_native_units():
    vs, ds = (9, 8, 7), (11, 6)
    vol = torch.from_numpy(np.random.default_rng(2).standard_normal((1, 1, 7, 8, 9)))
    P = G.beijing_default_vectors(3, 4, ds, 30.0, 50.0, (0.25, 0.3))
    g = G.beijing(P, vs, ds)
    a = cone.bj_forward(vol, P, vs, ds)
    assert float((a - _table_forward(g, vol)).abs().max()) < 1e-12


def test_circular_table_matches_native_units():
    vs, ds = (9, 8, 7), (11, 6)
    ang = np.linspace(0, 360, 8, False) + 1.5
    vol = torch.from_numpy(np.random.default_rng(3).standard_normal((1, 1, 7, 8, 9)))
    g = G.circular(ang, vs, 0.7, ds, (0.9, 1.1), 20.0, 35.0)
    a = cone.circ_forward(vol, ang, vs, 0.7, ds, (0.9, 1.1), 20.0, 35.0)
    assert float((a - _table_forward(g, vol)).abs().max()) < 1e-12


def test_subset_views():
    g = configs.fd_32()
    s = g.subset([3, 5])
    assert s.n_views == 2 and np.array_equal(s.views, g.views[[3, 5]])


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "dtvproj.h")).read()
    return sorted(set(re.findall(r"\b(dtv_[a-z_0-9]+)\s*\(", txt)))


def test_library_loads_and_exports_header_symbols():
    lib = _capi.lib()
    syms = _header_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_capi.EXPORTED)
    assert "gfx950" in _capi.version()


def test_library_is_built_from_the_sources_beside_it():
    """The loaded library carries the digest of the sources it was compiled
    from, and the loader refuses one whose digest differs (a stale build)."""
    from dtvproj import build
    want = build.source_digest()
    assert want is not None and len(want) == 16
    assert _capi.version().endswith("src:" + want)
    _capi.check_fresh(_capi.version())
    flipped = want[:-1] + ("0" if want[-1] != "0" else "1")
    for stale in ("dtvproj 0.2.0 gfx950 src:" + flipped, "dtvproj 0.1.0 gfx950",
                  "dtvproj 0.2.0 gfx950 src:unknown"):
        with pytest.raises(RuntimeError, match="rebuild"):
            _capi.check_fresh(stale)
    # the digest covers every source and header the library is compiled from
    names = {os.path.basename(f) for f in build.digest_inputs()}
    assert set(build.SOURCES) <= names and "dtvproj.h" in names and "dtv_geom.h" in names


def test_host_view_builders_match_python_geometry():
    vs, ds = (37, 29, 21), (45, 27)
    ang = np.linspace(0, 360, 13, False) + 0.3
    views, sc = ops.host_views_standard(ang, vs, ds, 700.0, 1300.0)
    g = G.standard(ang, vs, ds, 700.0, 1300.0)
    np.testing.assert_allclose(views, g.views, rtol=0, atol=1e-9)
    assert sc.axis_mode == g.scheme.axis_mode and sc.margin == pytest.approx(g.scheme.margin)
    assert sc.scale == pytest.approx(g.scheme.scale)
    assert list(sc.box_lo) == pytest.approx(list(g.scheme.box_lo))
    P = G.beijing_default_vectors(2, 21, ds, 300.0, 500.0, (0.5, 0.6))
    views, sc = ops.host_views_beijing(P, vs, ds)
    gb = G.beijing(P, vs, ds)
    np.testing.assert_allclose(views, gb.views, rtol=0, atol=1e-9)
    assert list(sc.box_hi) == list(gb.scheme.box_hi)


def test_capi_rejects_bad_arguments_without_gpu():
    lib = _capi.lib()
    out = np.zeros((1, 12))
    sc = _capi.DtvScheme()
    a = np.zeros(1)
    cp = a.ctypes.data_as(_capi.c_double_p)
    op = out.ctypes.data_as(_capi.c_double_p)
    # zero-sized volume, sdd <= sid, null outputs
    assert lib.dtv_views_standard(cp, 1, _capi.int3((0, 4, 4)), _capi.int2((4, 4)), 1.0, 2.0, op,
                                  ctypes.byref(sc)) == 1
    assert lib.dtv_views_standard(cp, 1, _capi.int3((4, 4, 4)), _capi.int2((4, 4)), 2.0, 1.0, op,
                                  ctypes.byref(sc)) == 1
    assert lib.dtv_views_standard(cp, 1, _capi.int3((4, 4, 4)), _capi.int2((4, 4)), 1.0, 2.0, None,
                                  ctypes.byref(sc)) == 2
    # the GPU entry points validate before touching the device
    assert lib.dtv_cone_forward(None, None, 1, _capi.int3((4, 4, 4)), _capi.int2((4, 4)), None, 1,
                                ctypes.byref(sc), None) == 2
    assert lib.dtv_beijing_backward(ctypes.c_void_p(16), ctypes.c_void_p(16), 1,
                                    _capi.int3((4, 4, 4)), _capi.int2((4, 4)),
                                    ctypes.c_void_p(16), 20, -1, None) == 5
    assert "multiple of 21" in lib.dtv_strerror(5).decode()
    # beyond 32-bit per-volume indexing (2048 x 2048 x 1024 voxels)
    sc.axis_mode = 3
    views = np.zeros((4, 12))
    assert lib.dtv_cone_forward(ctypes.c_void_p(16), ctypes.c_void_p(16), 1,
                                _capi.int3((2048, 2048, 1024)), _capi.int2((64, 64)),
                                views.ctypes.data_as(_capi.c_double_p), 4, ctypes.byref(sc),
                                None) == 3
    # view splits beyond what the kernels take (A: 16 parts, A^T: 2)
    views4 = np.zeros((4, 12))
    vp = views4.ctypes.data_as(_capi.c_double_p)
    for fn, parts in ((lib.dtv_cone_forward_ex, 17), (lib.dtv_cone_forward_ex, -1),
                      (lib.dtv_cone_adjoint_ex, 3)):
        L = _capi.launch_struct(algorithm=2, brick=(16, 16, 8), block=256, view_parts=parts)
        assert fn(ctypes.c_void_p(16), ctypes.c_void_p(16), 1, _capi.int3((8, 8, 8)),
                  _capi.int2((8, 8)), vp, 4, ctypes.byref(sc), ctypes.byref(L), None) == 1
    # the split adjoint: a workspace is needed beyond one part, and big enough
    L = _capi.launch_struct(algorithm=2, brick=(16, 16, 8), block=256, view_parts=3)
    args = (ctypes.c_void_p(16), ctypes.c_void_p(16), 1, _capi.int3((8, 8, 8)), _capi.int2((8, 8)),
            vp, 4, ctypes.byref(sc), ctypes.byref(L))
    assert lib.dtv_cone_adjoint_ws(*args, None, 0, None) == 2
    need = _part_ws_bytes((8, 8, 8), (16, 16, 8), 1, 3)
    assert lib.dtv_cone_adjoint_ws(*args, ctypes.c_void_p(64), need - 1, None) == 1
    L.view_parts = 17
    assert lib.dtv_cone_adjoint_ws(*args, ctypes.c_void_p(64), 1 << 20, None) == 1
    # the deterministic forward: a workspace of 256 + 8 bytes per ray, 256-byte aligned
    L = _capi.launch_struct(algorithm=2, brick=(16, 16, 8), block=256)
    fargs = (ctypes.c_void_p(16), ctypes.c_void_p(16), 1, _capi.int3((8, 8, 8)), _capi.int2((8, 8)),
             vp, 4, ctypes.byref(sc), ctypes.byref(L))
    need = 256 + 8 * 4 * 8 * 8
    assert lib.dtv_cone_forward_ws(*fargs, None, 0, None) == 2
    assert lib.dtv_cone_forward_ws(*fargs, ctypes.c_void_p(256), need - 1, None) == 1
    assert lib.dtv_cone_forward_ws(*fargs, ctypes.c_void_p(128), need, None) == 1
    # TV steps and the row filter: null pointers, aliasing, bad sizes and modes
    p = ctypes.c_void_p(16)
    q = ctypes.c_void_p(32)
    v3 = _capi.int3((4, 4, 4))
    f32 = ctypes.c_float
    assert lib.dtv_tv_descent(None, q, 1, v3, f32(0.1), f32(1.0), 1, None) == 2
    assert lib.dtv_tv_descent(p, p, 1, v3, f32(0.1), f32(1.0), 1, None) == 1   # in place
    assert lib.dtv_tv_descent(p, q, 1, v3, f32(0.1), f32(0.0), 1, None) == 1   # eps <= 0
    assert lib.dtv_tv_descent(p, q, 1, v3, f32(0.1), f32(1.0), 2, None) == 1   # mode
    assert lib.dtv_tv_descent(p, q, 0, v3, f32(0.1), f32(1.0), 0, None) == 1   # batch
    assert lib.dtv_tv_descent(p, q, 1, _capi.int3((4, 0, 4)), f32(0.1), f32(1.0), 0,
                              None) == 1
    r = ctypes.c_void_p(48)
    s = ctypes.c_void_p(64)
    lam = (f32 * 3)(0.1, 0.1, 0.1)
    neg = (f32 * 3)(0.1, -0.1, 0.1)
    assert lib.dtv_pdhg_dtv_step(p, q, r, None, 1, v3, f32(0.5), f32(0.5), lam, 0, None) == 2
    assert lib.dtv_pdhg_dtv_step(p, q, r, s, 1, v3, f32(0.5), f32(0.5), None, 0, None) == 2
    assert lib.dtv_pdhg_dtv_step(p, p, r, s, 1, v3, f32(0.5), f32(0.5), lam, 0, None) == 1
    assert lib.dtv_pdhg_dtv_step(p, q, q, s, 1, v3, f32(0.5), f32(0.5), lam, 0, None) == 1
    assert lib.dtv_pdhg_dtv_step(p, q, r, p, 1, v3, f32(0.5), f32(0.5), lam, 0, None) == 1
    assert lib.dtv_pdhg_dtv_step(p, q, r, s, 1, v3, f32(0.5), f32(0.5), neg, 0, None) == 1
    assert lib.dtv_pdhg_dtv_step(p, q, r, s, 1, v3, f32(0.5), f32(0.5), lam, 3, None) == 1
    assert lib.dtv_pdhg_dtv_step(p, q, r, s, 0, v3, f32(0.5), f32(0.5), lam, 0, None) == 1
    assert lib.dtv_row_filter(p, q, 4, 8, None, 0, p, 4, None) == 1            # even taps
    assert lib.dtv_row_filter(p, p, 4, 8, None, 0, q, 5, None) == 1            # in place
    assert lib.dtv_row_filter(None, q, 4, 8, None, 0, p, 5, None) == 2
    assert lib.dtv_row_filter(p, q, 4, 20000, None, 0, ctypes.c_void_p(48), 5001,
                              None) == 3                                       # LDS budget
    # FISTA-TV step: outputs may not overwrite y (stencil input) or each other
    one = f32(1.0)
    assert lib.dtv_tv_fista_step(p, q, r, p, s, 1, v3, one, one, one, one, 1, 0, None) == 1
    assert lib.dtv_tv_fista_step(p, q, r, s, p, 1, v3, one, one, one, one, 1, 0, None) == 1
    assert lib.dtv_tv_fista_step(p, q, r, s, s, 1, v3, one, one, one, one, 1, 0, None) == 1
    # shrink step: ynew may be neither xprev nor xnew; at least one value
    assert lib.dtv_shrink_step(p, q, r, s, r, 4, one, one, one, None) == 1
    assert lib.dtv_shrink_step(p, q, r, s, s, 4, one, one, one, None) == 1
    assert lib.dtv_shrink_step(p, q, r, s, None, 4, one, one, one, None) == 2
    assert lib.dtv_shrink_step(p, q, r, s, ctypes.c_void_p(80), 0, one, one, one, None) == 1


def test_ops_require_device_tensors():
    g = configs.fd_32()
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        ops.forward(torch.zeros(g.volume_shape(1)), g)
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        ops.adjoint(torch.zeros(g.sino_shape(1)), g)


def test_jit_bindings_reference_error_messages():
    import JITBeijingGeometry as JB
    import JITStandardGeometry as JS
    with pytest.raises(RuntimeError, match="volume must be a CUDA tensor"):
        JS.forward(torch.zeros(1, 1, 4, 4, 4), torch.zeros(3, dtype=torch.float64),
                   torch.tensor([4, 4, 4]), torch.tensor([4, 4]), 1.0, 2.0, 0)
    with pytest.raises(RuntimeError, match="sino must be a CUDA tensor"):
        JS.backward(torch.zeros(1, 1, 3, 4, 4), torch.zeros(3, dtype=torch.float64),
                    torch.tensor([4, 4, 4]), torch.tensor([4, 4]), 1.0, 2.0, 0)
    with pytest.raises(RuntimeError, match="volume must be a CUDA tensor"):
        JB.forward(torch.zeros(1, 1, 4, 4, 4), torch.tensor([4, 4, 4]), torch.tensor([4, 4]),
                   torch.zeros(21, 12), 0)


def test_sample_count_matches_oracle_enumeration():
    g = configs.fd_32()
    sc = g.scheme
    r = cone.table_rays(g.views, g.detector_size, sc.axis_mode, sc.box_lo, sc.box_hi, sc.margin,
                        sc.scale)
    assert roofline.sample_count(g) == cone.sample_count(r, g.volume_size)
    ang = G.uniform_angles(8)
    gs = G.standard(ang, (40, 36, 20), (50, 16), 100.0, 200.0)
    rs = cone.std_rays(ang, (40, 36, 20), (50, 16), 100.0, 200.0)
    assert roofline.sample_count(gs) == cone.sample_count(rs, (40, 36, 20))


def test_pair_bytes_accounting():
    g = configs.fd_32()
    b = roofline.pair_bytes(g, samples=1000)
    assert b["forward"] == 32 * 1000 + 4 * g.n_rays
    assert b["adjoint"] == 32 * 1000 + 4 * 32 ** 3
    assert b["pair"] == b["forward"] + b["adjoint"]


def _part_ws_bytes(vol, brick, batch, parts):
    """Workspace of the partitioned A^T: the partial volumes of parts 1..,
    then each part's upper-halo sums, (b+1)^3 - b^3 floats per brick (for
    b = (bx, by, bz) per axis)."""
    X, Y, Z = vol
    bx, by, bz = brick
    nb = -(-X // bx) * -(-Y // by) * -(-Z // bz)
    halo = (bx + 1) * (by + 1) * (bz + 1) - bx * by * bz
    return ((parts - 1) * batch * X * Y * Z + parts * batch * nb * halo) * 4


def test_adjoint_plan_splits_by_volume_size():
    """The workspace split the library picks (no GPU work): two parts at
    256^3 / 60 views, none at 512x512x384, many on small volumes (at most 16),
    and the workspace holds the partial volumes of parts 1.. and every part's
    upper-halo sums."""
    from dtvproj import configs
    cases = [(configs.sparse_256(), 1, (32, 16, 16), 2), (configs.large_512(), 1, (32, 16, 16), 1),
             (configs.fista_64(), 1, (16, 16, 16), 16), (configs.unroll_128(), 4, (32, 16, 16), 4)]
    for g, batch, brick, parts in cases:
        plan, nbytes = ops.adjoint_plan(g, batch)
        assert plan.algorithm == 2 and tuple(plan.brick) == brick and plan.view_parts == parts
        assert nbytes == _part_ws_bytes(g.volume_size, brick, batch, parts)
    # halo count of a 32x16x16 brick: the z = 16 plane 33 x 17, the y = 16
    # rows of the planes below 33 x 16, the x = 32 column of the rest 16 x 16
    assert _part_ws_bytes((32, 16, 16), (32, 16, 16), 1, 1) == (33 * 17 + 33 * 16 + 16 * 16) * 4
    # a fixed launch keeps its shape and split
    plan, nbytes = ops.adjoint_plan(configs.fista_64(), 1,
                                    dict(algorithm=2, brick=(8, 8, 8), block=64, view_parts=5))
    assert tuple(plan.brick) == (8, 8, 8) and plan.block == 64 and plan.view_parts == 5
    assert nbytes == _part_ws_bytes((64, 64, 64), (8, 8, 8), 1, 5)
    # the ray-driven kernel needs none
    plan, nbytes = ops.adjoint_plan(configs.fista_64(), 1, dict(algorithm=1))
    assert plan.algorithm == 1 and nbytes == 0


def test_forward_plan_workspace():
    """The deterministic forward's plan: the default brick launch and 256 +
    8 bytes per ray of workspace; none for the ray-driven kernel."""
    from dtvproj import configs
    g = configs.sparse_256()
    plan, nbytes = ops.forward_plan(g, 2)
    assert plan.algorithm == 2 and tuple(plan.brick) == (26, 26, 14) and plan.view_parts >= 1
    assert nbytes == 256 + 8 * 2 * 60 * 384 * 384
    plan, nbytes = ops.forward_plan(g, 1, dict(algorithm=1))
    assert plan.algorithm == 1 and nbytes == 0


def test_multi_adjoint_workspace_bytes_and_checks():
    """The partitioned multi-instance A^T's workspace: a 256-byte header per
    16 instances (per-instance offsets, computed on the device), then each
    instance's view parts x bricks x halo floats; the entry point refuses a
    missing, small or misaligned workspace before any GPU work."""
    from dtvproj.batched import Instance, WorkItem
    host = (Instance * 2)()
    shapes = [((45, 38, 29), (32, 16, 16), 2), ((20, 16, 12), (7, 5, 3), 1)]
    want = 256
    for e, (vs, b, parts) in zip(host, shapes):
        e.nx, e.ny, e.nz = vs
        e.nu, e.nv, e.n_views = 16, 16, 4
        e.axis_mode = 3
        e.bx, e.by, e.bz = b
        e.nbx, e.nby, e.nbz = (-(-vs[0] // b[0]), -(-vs[1] // b[1]), -(-vs[2] // b[2]))
        e.view_parts = parts
        e.in_ = e.out = e.views = 16
        halo = (b[0] + 1) * (b[1] + 1) * (b[2] + 1) - b[0] * b[1] * b[2]
        want += 4 * parts * e.nbx * e.nby * e.nbz * halo
    lib = _capi.lib()
    n = ctypes.c_size_t(0)
    assert lib.dtv_cone_adjoint_multi_ws_bytes(ctypes.addressof(host), 2, ctypes.byref(n)) == 0
    assert n.value == want
    assert lib.dtv_cone_adjoint_multi_ws_bytes(None, 2, ctypes.byref(n)) == 2
    assert lib.dtv_cone_adjoint_multi_ws_bytes(ctypes.addressof(host), 0, ctypes.byref(n)) == 1
    work = (WorkItem * 1)()
    args = (ctypes.addressof(host), ctypes.c_void_p(256), 2, ctypes.addressof(work),
            ctypes.c_void_p(256), 1, 256, None, 0)
    assert lib.dtv_cone_adjoint_multi_ws(*args, None, 0, None) == 2
    assert lib.dtv_cone_adjoint_multi_ws(*args, ctypes.c_void_p(256), want - 1, None) == 1
    assert lib.dtv_cone_adjoint_multi_ws(*args, ctypes.c_void_p(128), want, None) == 1
    host[0].view_parts = 3   # the multi-instance A^T splits a brick at most in two
    assert lib.dtv_cone_adjoint_multi_ws_bytes(ctypes.addressof(host), 2, ctypes.byref(n)) == 1
