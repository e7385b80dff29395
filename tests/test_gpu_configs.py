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
