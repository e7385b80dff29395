"""GPU parity over seeded random geometries: every kernel family against the
FP64 oracle on small circular and Standard-scheme problems whose sizes,
magnifications, detector shapes and view angles are drawn from fixed seeds
(odd and non-multiple-of-tile sizes, short and wide detectors, irregular
angles, sources close to the volume).  Bar: relative L2 <= 1e-5 for A, the
matched A^T and the reference backprojector against the oracle, and an adjoint
dot-product mismatch <= 1e-6 of ||Ax|| ||y|| (the inputs are N(0,1), whose
dot products nearly cancel).
"""
import numpy as np
import pytest
import torch

from conftest import rel_l2

from oracle import cone
from dtvproj import geometry as G, ops, phantom

pytestmark = pytest.mark.gpu

TOL = 1e-5
LAUNCHES = [None, dict(algorithm=1), dict(algorithm=2, brick=(6, 5, 7), block=128),
            dict(algorithm=2, brick=(16, 16, 8), block=256)]


def _circ_case(seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = (int(v) for v in rng.integers(5, 41, size=3))
    vx = float(rng.uniform(0.5, 1.5))
    U, V = (int(v) for v in rng.integers(8, 49, size=2))
    half_diag = 0.5 * vx * float(np.sqrt(X * X + Y * Y + Z * Z))
    # every fourth case puts the source just outside the volume's sphere
    sod = half_diag * (1.1 if seed % 4 == 0 else float(rng.uniform(1.5, 6.0)))
    sdd = sod * float(rng.uniform(1.2, 3.0))
    # the detector covers the volume's shadow, more or less
    cover = float(rng.uniform(0.6, 1.4))
    pitch = (2.0 * half_diag * sdd / sod * cover / U, 2.0 * half_diag * sdd / sod * cover / V)
    n = int(rng.integers(3, 17))
    ang = np.sort(rng.uniform(0.0, 360.0, size=n))
    return dict(vol=(X, Y, Z), vx=vx, det=(U, V), pitch=pitch, sod=sod, sdd=sdd, ang=ang)


@pytest.mark.parametrize("seed", range(24))
def test_random_circular_geometry_vs_oracle(cuda, seed):
    c = _circ_case(seed)
    X, Y, Z = c["vol"]
    U, V = c["det"]
    print(f"[case] random circular seed={seed} vol={c['vol']} det={c['det']} "
          f"views={len(c['ang'])} sod={c['sod']:.1f} sdd={c['sdd']:.1f}", flush=True)
    g = G.circular(c["ang"], c["vol"], c["vx"], c["det"], c["pitch"], c["sod"], c["sdd"])
    args = (c["ang"], c["vol"], c["vx"], c["det"], c["pitch"], c["sod"], c["sdd"])
    x = phantom.gaussian((2, 1, Z, Y, X), 300 + seed)
    y = phantom.gaussian((2, 1, len(c["ang"]), V, U), 400 + seed)
    want_a = cone.circ_forward(x, *args)
    want_t = cone.circ_adjoint(y, *args)
    xd, yd = x.to(cuda), y.to(cuda)
    for launch in LAUNCHES:
        a = ops.forward(xd, g, launch)
        t = ops.adjoint(yd, g, launch)
        assert rel_l2(a, want_a) <= TOL, (launch, rel_l2(a, want_a))
        assert rel_l2(t, want_t) <= TOL, (launch, rel_l2(t, want_t))
        # N(0,1) inputs make <Ax, y> a near-cancelling sum, so the mismatch is
        # measured against the Cauchy-Schwarz scale ||Ax|| ||y||
        lhs = float((a.double().cpu() * y.double()).sum())
        rhs = float((x.double() * t.double().cpu()).sum())
        scale = float(a.double().norm()) * float(y.double().norm())
        assert abs(lhs - rhs) <= 1e-6 * scale, launch


@pytest.mark.parametrize("seed", range(16))
def test_random_standard_reference_backprojector_vs_oracle(cuda, seed):
    """The Standard reference backprojector over random sizes, distances and
    short detectors, so that the 16-voxel z runs of its lanes start, end and
    straddle the detector's top and bottom rows and the volume is narrower
    than a wave."""
    rng = np.random.default_rng(1000 + seed)
    X, Y, Z = (int(v) for v in rng.integers(6, 70, size=3))
    U = int(rng.integers(16, 96))
    V = int(rng.integers(6, 40))
    sid = float(rng.uniform(1.2, 4.0)) * 0.5 * float(np.hypot(X, Y))
    sdd = sid * float(rng.uniform(1.3, 2.5))
    ang = np.sort(rng.uniform(0.0, 360.0, size=int(rng.integers(3, 13))))
    print(f"[case] random standard bp_ref seed={seed} vol={(X, Y, Z)} det={(U, V)} "
          f"views={len(ang)} sid={sid:.1f} sdd={sdd:.1f}", flush=True)
    g = G.standard(ang, (X, Y, Z), (U, V), sid, sdd)
    y = phantom.gaussian((1, 1, len(ang), V, U), 500 + seed)
    got = ops.bp_ref(y.to(cuda), g)
    want = cone.std_bp_ref(y, ang, (X, Y, Z), (U, V), sid, sdd)
    assert rel_l2(got, want) <= TOL


def _rot(axis, deg):
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    i, j = [(1, 2), (0, 2), (0, 1)][axis]
    m = np.eye(3)
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


@pytest.mark.parametrize("seed", range(10))
def test_random_beijing_scan_vs_oracle(cuda, seed):
    """Beijing view vectors (21 sub-detectors per source angle) with random
    volume and strip sizes, distances and pitches, each view's detector
    tilted by up to 3 degrees and shifted by up to a voxel, so no two views
    share an orientation: A and the matched A^T on every kernel family, and
    the reference backprojector."""
    rng = np.random.default_rng(2000 + seed)
    X, Y, Z = (int(v) for v in rng.integers(6, 41, size=3))
    U, V = int(rng.integers(3, 13)), int(rng.integers(6, 33))
    sid = float(rng.uniform(0.8, 4.0)) * 0.5 * float(np.hypot(X, Y))
    sdd = sid + float(rng.uniform(10.0, 60.0))
    pitch = (float(rng.uniform(0.4, 1.2)), float(rng.uniform(0.5, 2.0)))
    P = np.asarray(G.beijing_default_vectors(int(rng.integers(1, 3)), 21, (U, V), sid, sdd, pitch),
                   dtype=np.float64)
    for r in P:
        m = _rot(int(rng.integers(0, 3)), float(rng.uniform(-3.0, 3.0)))
        r[3:6] += rng.uniform(-1.0, 1.0, size=3)
        r[6:9] = m @ r[6:9]
        r[9:12] = m @ r[9:12]
    P = P.astype(np.float32)
    print(f"[case] random beijing seed={seed} vol={(X, Y, Z)} det={(U, V)} views={len(P)} "
          f"sid={sid:.1f} sdd={sdd:.1f}", flush=True)
    vs, ds = (X, Y, Z), (U, V)
    g = G.beijing(P, vs, ds)
    x = phantom.gaussian((1, 1, Z, Y, X), 600 + seed)
    y = phantom.gaussian(g.sino_shape(1), 700 + seed)
    want_a = cone.bj_forward(x, P, vs, ds)
    want_t = cone.bj_adjoint(y, P, vs, ds)
    for launch in LAUNCHES:
        a = ops.forward(x.to(cuda), g, launch)
        t = ops.adjoint(y.to(cuda), g, launch)
        assert rel_l2(a, want_a) <= TOL, (launch, rel_l2(a, want_a))
        assert rel_l2(t, want_t) <= TOL, (launch, rel_l2(t, want_t))
    assert rel_l2(ops.bp_ref(y.to(cuda), g), cone.bj_bp_ref(y, P, vs, ds)) <= TOL


@pytest.mark.parametrize("seed", range(6))
def test_random_multi_instance_launch_vs_oracle(cuda, seed):
    """dtv_cone_forward_multi / dtv_cone_adjoint_multi over 3-5 random
    instances in one launch (random circular geometries, at times a random
    Beijing scan among them), one shared brick shape per launch drawn from
    compiled and run-time shapes, grid and persistent workgroups alternating
    with the seed, each instance's output against the FP64 oracle."""
    from dtvproj import batched
    rng = np.random.default_rng(3000 + seed)
    n = int(rng.integers(3, 6))
    geoms, oracles = [], []
    for i in range(n):
        if i == n - 1 and seed % 2 == 0:
            X, Y, Z = (int(v) for v in rng.integers(8, 33, size=3))
            U, V = int(rng.integers(4, 12)), int(rng.integers(8, 25))
            sid = float(rng.uniform(1.5, 3.0)) * 0.5 * float(np.hypot(X, Y))
            P = G.beijing_default_vectors(int(rng.integers(1, 3)), 21, (U, V), sid,
                                          sid + float(rng.uniform(20.0, 50.0)),
                                          (float(rng.uniform(0.5, 1.0)), 1.0))
            geoms.append(G.beijing(P, (X, Y, Z), (U, V)))
            oracles.append((lambda P, vs, ds: (lambda x: cone.bj_forward(x, P, vs, ds),
                                               lambda y: cone.bj_adjoint(y, P, vs, ds)))(
                P, (X, Y, Z), (U, V)))
        else:
            c = _circ_case(3100 + 10 * seed + i)
            geoms.append(G.circular(c["ang"], c["vol"], c["vx"], c["det"], c["pitch"], c["sod"],
                                    c["sdd"]))
            args = (c["ang"], c["vol"], c["vx"], c["det"], c["pitch"], c["sod"], c["sdd"])
            oracles.append((lambda args: (lambda x: cone.circ_forward(x, *args),
                                          lambda y: cone.circ_adjoint(y, *args)))(args))
    shapes = [((26, 26, 14), (32, 16, 16)), ((16, 16, 8), (16, 16, 8)), ((6, 5, 7), (9, 7, 5)),
              ((24, 24, 14), (32, 16, 16))]
    fb, ab = shapes[seed % len(shapes)]
    persistent = 0 if seed % 2 else 23
    print(f"[case] random multi seed={seed} instances={n} bricks={fb},{ab} "
          f"persistent={persistent}", flush=True)
    mp = batched.MultiProjector(geoms, bricks=[(fb, ab)] * n, block=256,
                                persistent_blocks=persistent)
    vols = [phantom.gaussian(g.volume_shape(1), 800 + 10 * seed + i) for i, g in enumerate(geoms)]
    ys = [phantom.gaussian(g.sino_shape(1), 900 + 10 * seed + i) for i, g in enumerate(geoms)]
    sinos = mp.forward([v.to(cuda) for v in vols])
    outs = mp.adjoint([y.to(cuda) for y in ys])
    for i, (fwd, adj) in enumerate(oracles):
        assert rel_l2(sinos[i], fwd(vols[i].double())) <= TOL, f"forward {i}"
        assert rel_l2(outs[i], adj(ys[i].double())) <= TOL, f"adjoint {i}"


def _rand_shape(rng, square=False):
    B = int(rng.integers(1, 3))
    Z = int(rng.integers(1, 40))
    X = int(rng.choice([int(rng.integers(1, 90)), 4 * int(rng.integers(1, 80))]))
    Y = X if square else int(rng.integers(1, 90))
    return (B, 1, Z, Y, X)


@pytest.mark.parametrize("seed", range(12))
def test_random_fused_volume_steps(cuda, seed):
    """The fused FISTA-TV step, the ASD-POCS descent step and the PDHG/DTV
    pair on random shapes (1 to 39 slices, widths with and without the
    four-voxel march, square and not, so reference mode takes both the march
    and the per-voxel kernels), both boundary modes, against their FP64
    formulations whose difference operators meet the sparse matrices."""
    from dtvproj import tv
    from oracle import tv as otv
    rng = np.random.default_rng(4000 + seed)
    mode = ("neumann", "reference")[seed % 2]
    shp = _rand_shape(rng, square=bool(seed % 3 == 0))
    Z, Y, X = shp[-3:]
    print(f"[case] random volume steps seed={seed} {mode} {shp}", flush=True)
    y, g, xp = (phantom.gaussian(shp, 40 * seed + k) for k in range(3))
    args = dict(step=0.3, lam=0.7, eps=0.05, beta=0.4, nonneg=True, mode=mode)
    rx, ry = tv.fista_tv_step_torch(y.double(), g.double(), xp.double(), **args)
    x, yn = tv.fista_tv_step(y.to(cuda), g.to(cuda), xp.to(cuda), **args)
    assert rel_l2(x, rx) <= 1e-6 and rel_l2(yn, ry) <= 1e-6
    want = y.double().numpy() - 0.25 * otv.asd_tv_grad(y.numpy(), 1.0, (X, Y, Z), mode)
    assert rel_l2(tv.tv_descent(y.to(cuda), 0.25, 1.0, mode), torch.from_numpy(want)) <= 1e-6
    f, fbar, atw = y - 0.3, g, xp
    p = torch.stack([phantom.gaussian(shp, 40 * seed + 10 + a) - 0.4 for a in range(3)])
    lam = (0.3, 0.5, 0.2)
    rf, rfb, rp = tv.pdhg_dtv_step_torch(f.double(), fbar.double(), atw.double(), p.double(),
                                         0.7, 0.4, lam, mode)
    fd, fbd, pd = f.to(cuda), fbar.to(cuda), p.to(cuda)
    tv.pdhg_dtv_step(fd, fbd, atw.to(cuda), pd, 0.7, 0.4, lam, mode)
    assert rel_l2(fd, rf) <= 1e-6 and rel_l2(fbd, rfb) <= 1e-6
    for a in range(3):
        assert rel_l2(pd[a], rp[a]) <= 1e-6
