"""GPU: multi-instance launches vs per-instance calls, the batched FISTA-TV
unroll (eager and hipGraph replay) vs the plain per-volume loop, autotuner."""
import pytest
import torch

from conftest import rel_l2

from oracle import cone
from dtvproj import autotune, batched, configs, geometry as G, ops, phantom, unroll

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _small_uneven():
    return [
        G.circular(G.uniform_angles(11), (40, 40, 24), 1.0, (60, 30), 1.6, 200.0, 400.0),
        G.circular(G.uniform_angles(7) + 3.0, (29, 29, 37), 1.1, (45, 52), 1.5, 150.0, 330.0),
        G.circular(G.uniform_angles(16), (52, 52, 17), 0.9, (80, 24), 1.7, 240.0, 500.0),
    ]


BRICKS = [((32, 16, 16), (32, 16, 16)), ((16, 16, 8), (8, 8, 8)), ((7, 5, 3), (16, 16, 16))]


@pytest.mark.parametrize("partition", [True, False])
@pytest.mark.parametrize("persistent", [0, 37])
def test_multi_matches_single_instance_calls(cuda, persistent, partition):
    """Each instance of a multi launch equals its single call with the same
    brick shape (A to FP32 rounding; A^T, partitioned or not, bitwise)."""
    geoms = _small_uneven()
    _case("multi_vs_single", persistent=persistent, partition=partition)
    mp = batched.MultiProjector(geoms, bricks=BRICKS, block=256, persistent_blocks=persistent,
                                view_parts=1, partition=partition)
    vols = [phantom.gaussian(g.volume_shape(1), 40 + i).to(cuda) for i, g in enumerate(geoms)]
    inits = [phantom.gaussian(g.sino_shape(1), 50 + i).to(cuda) for i, g in enumerate(geoms)]
    sinos = mp.forward(vols, inits)
    for i, g in enumerate(geoms):
        ref = ops.forward(vols[i], g, dict(algorithm=2, brick=BRICKS[i][0], block=256)) - inits[i]
        assert rel_l2(sinos[i], ref) <= 1e-6
    ys = [phantom.gaussian(g.sino_shape(1), 60 + i).to(cuda) for i, g in enumerate(geoms)]
    outs = mp.adjoint(ys)
    for i, g in enumerate(geoms):
        # one workgroup per brick, as in the multi-instance launch
        ref = ops.adjoint(ys[i], g, dict(algorithm=2, brick=BRICKS[i][1], block=256,
                                         view_parts=1), partition=partition)
        assert torch.equal(outs[i], ref)   # integer accumulation: bitwise equal


@pytest.mark.parametrize("partition", [True, False])
@pytest.mark.parametrize("persistent", [0, 37])
def test_multi_with_a_culled_many_view_instance(cuda, persistent, partition):
    """A 147-view Beijing instance (more than four chunks: each brick culls the
    views it cannot reach) beside a small circular one, in one launch and in
    persistent workgroups that take several (instance, brick) items in turn:
    equal to the single-instance calls."""
    P = G.beijing_default_vectors(n_angles=7, planes=21, sub_detector=(8, 24), sid=70.0,
                                  sdd=114.0)
    geoms = [G.beijing(P, (40, 36, 16), (8, 24)),
             G.circular(G.uniform_angles(11), (40, 40, 24), 1.0, (60, 30), 1.6, 200.0, 400.0)]
    bricks = [((16, 16, 8), (16, 16, 8)), ((24, 24, 14), (32, 16, 16))]
    _case("multi_culled", persistent=persistent, views=[g.n_views for g in geoms],
          partition=partition)
    mp = batched.MultiProjector(geoms, bricks=bricks, block=256, persistent_blocks=persistent,
                                view_parts=1, partition=partition)
    vols = [phantom.gaussian(g.volume_shape(1), 70 + i).to(cuda) for i, g in enumerate(geoms)]
    sinos = mp.forward(vols)
    for i, g in enumerate(geoms):
        ref = ops.forward(vols[i], g, dict(algorithm=2, brick=bricks[i][0], block=256))
        assert rel_l2(sinos[i], ref) <= 1e-6
    ys = [phantom.gaussian(g.sino_shape(1), 80 + i).to(cuda) for i, g in enumerate(geoms)]
    outs = mp.adjoint(ys)
    for i, g in enumerate(geoms):
        ref = ops.adjoint(ys[i], g, dict(algorithm=2, brick=bricks[i][1], block=256,
                                         view_parts=1), partition=partition)
        assert torch.equal(outs[i], ref)


def _oracle_pair(g):
    """FP64 oracle A and A^T from the geometry's reference-form parameters
    (circular: angles / mm sizes; Beijing: the float32 view vectors)."""
    r = g.ref
    if "vectors" in r:
        args = (r["vectors"], g.volume_size, g.detector_size)
        return (lambda x: cone.bj_forward(x, *args)), (lambda y: cone.bj_adjoint(y, *args))
    args = (r["angles"], g.volume_size, r["voxel_mm"], g.detector_size, r["pitch_mm"], r["sod"],
            r["sdd"])
    return (lambda x: cone.circ_forward(x, *args)), (lambda y: cone.circ_adjoint(y, *args))


@pytest.mark.parametrize("partition", [True, False])
@pytest.mark.parametrize("persistent", [0, 37])
def test_multi_instance_kernels_vs_fp64_oracle(cuda, persistent, partition):
    """dtv_cone_forward_multi / dtv_cone_adjoint_multi on four uneven
    instances (three circular orbits of different size, voxel, pitch and view
    count, and a 147-view Beijing scan that the bricks cull), one workgroup
    per item and persistent workgroups, each output against the FP64 oracle
    (rel L2 <= 1e-5), the forward also with an initial value subtracted."""
    P = G.beijing_default_vectors(n_angles=7, planes=21, sub_detector=(8, 24), sid=70.0,
                                  sdd=114.0)
    geoms = _small_uneven() + [G.beijing(P, (40, 36, 16), (8, 24))]
    bricks = [((24, 24, 14), (32, 16, 16)), ((16, 16, 8), (8, 8, 8)), ((7, 5, 3), (16, 16, 16)),
              ((16, 16, 8), (16, 16, 8))]
    _case("multi vs oracle", persistent=persistent, views=[g.n_views for g in geoms])
    # bricks split over several items (A up to 5 parts, A^T 2) beside whole ones
    mp = batched.MultiProjector(geoms, bricks=bricks, block=256, persistent_blocks=persistent,
                                view_parts=[(3, 2), (1, 1), (5, 2), (4, 2)], partition=partition)
    vols = [phantom.gaussian(g.volume_shape(1), 140 + i) for i, g in enumerate(geoms)]
    inits = [phantom.gaussian(g.sino_shape(1), 150 + i) for i, g in enumerate(geoms)]
    ys = [phantom.gaussian(g.sino_shape(1), 160 + i) for i, g in enumerate(geoms)]
    sinos = mp.forward([v.to(cuda) for v in vols])
    resid = mp.forward([v.to(cuda) for v in vols], [s.to(cuda) for s in inits])
    outs = mp.adjoint([y.to(cuda) for y in ys])
    for i, g in enumerate(geoms):
        A64, AT64 = _oracle_pair(g)
        ref = A64(vols[i].double())
        assert rel_l2(sinos[i], ref) <= 1e-5, f"forward {i}"
        assert rel_l2(resid[i], ref - inits[i].double()) <= 1e-5, f"forward-init {i}"
        assert rel_l2(outs[i], AT64(ys[i].double())) <= 1e-5, f"adjoint {i}"


def test_batched_fista_tv_eager_graph_and_plain(cuda):
    geoms = _small_uneven()
    _case("batched_fista_tv")
    b = []
    for i, g in enumerate(geoms):
        ph = phantom.phantom(g.volume_size, 1.0, 4, seed=70 + i).to(cuda)
        b.append(ops.forward(ph, g))
    lam, eps, iters = 1e-3, 1e-3, 8
    steps = [1.0 / (unroll.power_norm(g, device=cuda) + lam * 12.0 / eps) for g in geoms]
    plain = batched.plain_fista_tv(geoms, b, iters, lam, eps, steps)
    bf = batched.BatchedFistaTV(geoms, b, iters, lam, eps, steps, bricks=BRICKS)
    bf.reset()
    eager = [x.clone() for x in bf.run()]
    for e, p in zip(eager, plain):
        assert rel_l2(e, p) <= 1e-4
    bf.capture()
    for rep in range(3):
        torch.cuda.synchronize()
        junk = torch.randn(1000, device=cuda).sum()   # eager work between replays
        out = bf.replay()
        torch.cuda.synchronize()
        for e, o in zip(eager, out):
            assert rel_l2(o, e) <= 1e-6, f"replay {rep}"
        float(junk)


def test_autotune_choices_are_valid_and_equivalent(cuda):
    g = configs.fd_32()
    _case("autotune fd_32")
    t = autotune.tune(g, cuda, reps=1)
    x = phantom.gaussian(g.volume_shape(1), 80).to(cuda)
    y = phantom.gaussian(g.sino_shape(1), 81).to(cuda)
    assert rel_l2(ops.forward(x, g, t["forward"]), ops.forward(x, g)) <= 1e-6
    assert rel_l2(ops.adjoint(y, g, t["adjoint"]), ops.adjoint(y, g)) <= 5e-6
    assert autotune.tune(g, cuda) is t   # cached


def test_shared_shape_tuning_for_multi_launches(cuda):
    """The multi-launch tuner returns one A shape and one A^T shape among its
    candidates, and a projector built with them matches the default one."""
    geoms = _small_uneven()
    _case("tune_multi")
    t = autotune.tune_multi(geoms, cuda, reps=1)
    shapes = {tuple(c[:3]) for c in autotune.CANDIDATES if c[3] == 256}
    assert t["forward"] in shapes and t["adjoint"] in shapes
    mp = batched.MultiProjector(geoms, bricks="tune_shared")
    assert mp.tuned is not None and mp.bricks[0] == (mp.tuned["forward"], mp.tuned["adjoint"])
    ref = batched.MultiProjector(geoms)
    vols = [phantom.gaussian(g.volume_shape(1), 90 + i).to(cuda) for i, g in enumerate(geoms)]
    for a, b in zip(mp.forward(vols), ref.forward(vols)):
        assert rel_l2(a, b) <= 1e-6
    ys = [phantom.gaussian(g.sino_shape(1), 95 + i).to(cuda) for i, g in enumerate(geoms)]
    for a, b in zip(mp.adjoint(ys), ref.adjoint(ys)):
        assert rel_l2(a, b) <= 5e-6


def test_partitioned_multi_adjoint_many_instances(cuda):
    """300 small uneven instances in one partitioned multi-instance A^T: the
    device-side offset scan then gives each of its lanes a run of two
    instances, and the face kernel finds each workgroup's instance by
    bisection over 300 prefixes.  Every output equals its single partitioned
    call with the same brick and split, bitwise."""
    import random
    rng = random.Random(7)
    geoms, bricks, parts = [], [], []
    for i in range(300):
        x, y, z = rng.randint(9, 30), rng.randint(9, 30), rng.randint(5, 20)
        geoms.append(G.circular(G.uniform_angles(rng.randint(3, 9)) + i, (x, y, z), 1.0,
                                (rng.randint(12, 40), rng.randint(8, 30)), 1.5, 60.0, 120.0))
        bricks.append(((16, 16, 8), (8, 8, 8) if i % 3 else (16, 16, 8)))
        parts.append((1, 2 if i % 4 == 0 else 1))
    _case("partitioned multi, 300 instances")
    mp = batched.MultiProjector(geoms, bricks=bricks, block=256, persistent_blocks=0,
                                view_parts=parts)
    ys = [phantom.gaussian(g.sino_shape(1), 900 + i).to(cuda) for i, g in enumerate(geoms)]
    outs = mp.adjoint(ys)
    for i, g in enumerate(geoms):
        ref = ops.adjoint(ys[i], g, dict(algorithm=2, brick=bricks[i][1], block=256,
                                         view_parts=parts[i][1]))
        assert torch.equal(outs[i], ref), f"instance {i}"
