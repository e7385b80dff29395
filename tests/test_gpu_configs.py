"""GPU: the two largest workloads, checked through size-independent
properties (the FP64 oracle would take hours at these sizes).

* 512x512x384 at 0.4 mm, 120 views, 768x576 at 0.6 mm (40-ellipsoid
  phantom, Poisson I0 = 5e4): adjoint dot-product identity, linearity, the
  brick kernels against the ray-driven ones at full size; a 30-iteration
  directional-TV (PDHG) reconstruction whose error to the phantom falls; a
  training-mode forward and backward through 5 checkpointed iterations.
* batch 4 of 128^3, 90 views, 192^2 (Poisson I0 = 1e5): one training step of
  the checkpointed 10-iteration PDHG/DTV network -- batch independence of the
  forward, finite gradients, and a loss that an SGD step lowers.
* the reference's own geometries at full size (Standard options: 256x256x64,
  48 views of 288x28; a 3024-view Beijing scan): A on a few views against
  the oracle, the adjoint identity, the reference backprojector's linearity.
"""
import pytest
import torch

from conftest import rel_l2
from dtvproj import configs, ops, phantom

pytestmark = pytest.mark.gpu


def _dot(a, b):
    return float((a.double() * b.double()).sum())


def test_large_512_dot_linearity_and_kernel_families(cuda):
    g = configs.large_512()
    print(f"[case] large_512 views={g.n_views} vol={g.volume_size} det={g.detector_size}", flush=True)
    x = phantom.phantom(g.volume_size, 0.4, 40, seed=5).to(cuda)
    y = ops.forward(x, g)
    # noisy measurement of the phantom at I0 = 5e4 as the A^T input
    yn = phantom.poisson_noise(y, 5e4, seed=6)
    z = ops.adjoint(yn, g)
    lhs, rhs = _dot(y, yn), _dot(x, z)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    # linearity of A on x + 2 u (u: N(0,1))
    u = phantom.gaussian(g.volume_shape(1), 7).to(cuda)
    assert rel_l2(ops.forward(x + 2 * u, g), y + 2 * ops.forward(u, g)) <= 1e-5
    del u
    # brick kernels vs the one-lane-per-ray kernels at full size
    ray = dict(algorithm=1)
    assert rel_l2(ops.forward(x, g, ray), y) <= 1e-6
    assert rel_l2(ops.adjoint(yn, g, ray), z) <= 5e-6


def test_unroll_128_batch4_training_step(cuda):
    from dtvproj import unroll
    g = configs.unroll_128()
    print(f"[case] unroll_128 batch 4 views={g.n_views}", flush=True)
    truth = torch.cat([phantom.phantom(g.volume_size, 1.0, 10, seed=30 + i) for i in range(4)])
    truth = (truth * 0.02).to(cuda)  # attenuation per mm
    with torch.no_grad():
        y = ops.forward(truth, g)
    y = phantom.poisson_noise(y, 1e5, seed=31)
    torch.manual_seed(0)
    net = unroll.PDHGDTVNet(g, iters=10, norm=unroll.power_norm(g, iters=6, device=cuda),
                            checkpointing=True).to(cuda)
    out = net(y)
    # batch independence of the whole network
    with torch.no_grad():
        one = net(y[2:3])
    assert rel_l2(out.detach()[2:3], one) <= 1e-5
    loss = torch.nn.functional.mse_loss(out, truth)
    loss.backward()
    grads = [p.grad for p in net.parameters()]
    assert all(gr is not None and torch.isfinite(gr).all() for gr in grads)
    assert any(float(gr.abs().max()) > 0 for gr in grads)
    with torch.no_grad():
        for p in net.parameters():
            p -= 1e-2 * p.grad / (p.grad.norm() + 1e-30)
        loss2 = torch.nn.functional.mse_loss(net(y), truth)
    assert float(loss2) < float(loss)


@pytest.fixture(scope="module")
def large_case(cuda):
    g = configs.large_512()
    truth = (phantom.phantom(g.volume_size, 0.4, 40, seed=5) * 0.02).to(cuda)
    with torch.no_grad():
        y = phantom.poisson_noise(ops.forward(truth, g), 5e4, seed=8)
    from dtvproj import unroll
    norm = unroll.power_norm(g, iters=6, device=cuda)
    return g, truth, y, norm


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def test_large_512_dtv_reconstruction(cuda, large_case):
    import time
    from dtvproj import unroll
    g, truth, y, norm = large_case
    print("[case] large_512 30-iteration DTV reconstruction", flush=True)
    lam = (2e-3, 2e-3, 4e-3)  # separate x / y / z weights
    errs = {}
    with torch.no_grad():
        for iters in (5, 30):
            net = unroll.PDHGDTVNet(g, iters=iters, lam=lam, norm=norm).to(cuda)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            f = net(y)
            torch.cuda.synchronize()
            errs[iters] = _rel(f, truth)
            print(f"  {iters} iterations: {time.perf_counter() - t0:.2f} s, "
                  f"rel error {errs[iters]:.4f}", flush=True)
            assert torch.isfinite(f).all()
            if iters == 5:
                # the fused volume updates vs the torch ops of the same network
                net.fused = False
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                plain = net(y)
                torch.cuda.synchronize()
                print(f"  5 iterations, torch volume updates: {time.perf_counter() - t0:.2f} s",
                      flush=True)
                assert _rel(f, plain) <= 1e-5
                del plain
    assert errs[30] < errs[5] < 1.0


def test_large_512_training_step_five_checkpointed_iterations(cuda, large_case):
    import time
    from dtvproj import unroll
    g, truth, y, norm = large_case
    print("[case] large_512 training step, 5 checkpointed iterations", flush=True)
    net = unroll.PDHGDTVNet(g, iters=5, lam=(2e-3, 2e-3, 4e-3), norm=norm,
                            checkpointing=True).to(cuda)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats(cuda)
    t0 = time.perf_counter()
    loss = torch.nn.functional.mse_loss(net(y), truth)
    loss.backward()
    torch.cuda.synchronize()
    peak = torch.cuda.max_memory_allocated(cuda) / 2 ** 30
    print(f"  forward + backward {time.perf_counter() - t0:.2f} s, peak {peak:.1f} GiB", flush=True)
    grads = [p.grad for p in net.parameters()]
    assert all(gr is not None and torch.isfinite(gr).all() for gr in grads)
    assert any(float(gr.abs().max()) > 0 for gr in grads)
    # checkpointing keeps one iteration's activations: well under one set
    # of all five iterations' intermediates
    assert peak < 40.0


# ---- the reference's own two geometries at full size ---------------------------

def _ref_configs():
    import numpy as np
    from dtvproj import geometry as G
    std = G.standard(G.uniform_angles(48), (256, 256, 64), (288, 28), 1000.0, 2000.0)
    X, Y, Z, U, V = 256, 256, 64, 80, 128
    mag = 1143.0 / 708.0
    pitch = (1.2 * np.hypot(X, Y) * mag / (U * 21), 1.2 * Z * mag / V)
    vec = G.beijing_default_vectors(144, 21, (U, V), 708.0, 1143.0, pitch)
    bj = G.beijing(vec, (X, Y, Z), (U, V))
    return std, bj, vec


def test_reference_geometries_full_size(cuda):
    """The Standard options geometry (256x256x64, 48 views of 288x28) and a
    3024-view Beijing scan of the same volume: A on a few views against the
    FP64 oracle (views are independent), the adjoint dot-product identity
    over all views, and the reference backprojector's linearity and its
    values at 300 voxels against the oracle."""
    import numpy as np
    from oracle import cone
    from dtvproj import geometry as G
    std, bj, vec = _ref_configs()
    x = phantom.phantom((256, 256, 64), 1.0, 20, seed=11)
    xd = x.to(cuda)
    for name, g in (("standard", std), ("beijing", bj)):
        print(f"[case] reference geometry {name} views={g.n_views}", flush=True)
        y = ops.forward(xd, g)
        pick = [0, g.n_views // 3, g.n_views - 1]
        if name == "standard":
            ref = cone.std_forward(x.double(), G.uniform_angles(48)[pick], (256, 256, 64),
                                   (288, 28), 1000.0, 2000.0)
        else:
            ref = cone.bj_forward(x.double(), vec[pick], (256, 256, 64), (80, 128))
        assert rel_l2(y[:, :, pick], ref) <= 1e-5
        # a noisy measurement of x as the A^T input (as in the 512 case): with
        # a pure N(0,1) sinogram and a non-negative x the dot product is a small
        # difference of large sums, below the FP32 rounding of either side
        s = y + phantom.gaussian(g.sino_shape(1), 12).to(cuda)
        z = ops.adjoint(s, g)
        lhs, rhs = _dot(y, s), _dot(xd, z)
        assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
        b1, b2 = ops.bp_ref(s, g), ops.bp_ref(y, g)
        assert rel_l2(ops.bp_ref(s + 2 * y, g), b1 + 2 * b2) <= 1e-5
        # the reference backprojector of the projection against the oracle at
        # 300 seeded voxels (all views)
        rng = np.random.default_rng(13)
        vox = np.stack([rng.integers(0, 256, 300), rng.integers(0, 256, 300),
                        rng.integers(0, 64, 300)], 1)
        yc = y.cpu()
        if name == "standard":
            want = cone.std_bp_ref(yc, G.uniform_angles(48), (256, 256, 64), (288, 28), 1000.0,
                                   2000.0, voxels=vox)
        else:
            want = cone.bj_bp_ref(yc, vec, (256, 256, 64), (80, 128), voxels=vox)
        got = b2[0, 0].cpu()[vox[:, 2], vox[:, 1], vox[:, 0]]
        assert rel_l2(got, want[0]) <= 1e-5
