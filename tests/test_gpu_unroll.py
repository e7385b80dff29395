"""GPU: autograd pair vs FP64 finite differences; FISTA-TV, the learned
PDHG/DTV unroll, ASD-POCS and the DTV Chambolle-Pock solver vs the same
algorithms driven by the FP64 oracle."""
import pytest
import torch

from conftest import rel_l2

from oracle import cone
from dtvproj import autograd, configs, ops, phantom, tv, unroll

pytestmark = pytest.mark.gpu


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


def _oracle_ops(g):
    """FP64 CPU A and A^T of a table geometry as differentiable functions."""
    sc = g.scheme
    args = (g.views, g.volume_size, g.detector_size, sc.axis_mode, sc.box_lo, sc.box_hi,
            sc.margin, sc.scale)

    class A(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x):
            return cone.table_forward(x, *args)

        @staticmethod
        def backward(ctx, gr):
            return cone.table_adjoint(gr, *args)

    class AT(torch.autograd.Function):
        @staticmethod
        def forward(ctx, y):
            return cone.table_adjoint(y, *args)

        @staticmethod
        def backward(ctx, gr):
            return cone.table_forward(gr, *args)

    return A.apply, AT.apply


def test_projector_gradients_vs_fp64_finite_differences(cuda):
    """d/dx 1/2 ||A x - b||^2 through the autograd Function (backward = the
    A^T kernel) vs the FP64 oracle gradient and FP64 central differences."""
    g = configs.fd_32()
    _case("fd_32 gradients")
    A64, AT64 = _oracle_ops(g)
    x = (phantom.phantom((32, 32, 32), 1.0, 6, seed=5) + 0.05 * phantom.gaussian(
        (1, 1, 32, 32, 32), 6)).double()
    # an O(1) residual: the FP32 forward's rounding (~1e-7 of A x) must stay
    # small against A x - b for a 1e-5 gradient comparison to be meaningful
    b = A64(x.clone()) + phantom.gaussian(g.sino_shape(1), 7, dtype=torch.float64)
    xg = x.float().to(cuda).requires_grad_(True)
    loss = 0.5 * ((autograd.project(xg, g) - b.float().to(cuda)) ** 2).sum()
    loss.backward()
    grad = xg.grad.double().cpu()
    ref = AT64(A64(x) - b)
    assert rel_l2(grad, ref) <= 1e-5
    L = lambda z: 0.5 * float(((A64(z) - b) ** 2).sum())
    for seed in (11, 12, 13):
        v = phantom.gaussian(x.shape, seed, dtype=torch.float64)
        h = 1e-3
        fd = (L(x + h * v) - L(x - h * v)) / (2 * h)
        an = float((grad * v).sum())
        assert abs(fd - an) <= 1e-5 * abs(fd)
    # and the backward of the backprojector is the A kernel
    yg = b.float().to(cuda).requires_grad_(True)
    c = phantom.gaussian(x.shape, 8).to(cuda)
    (0.5 * ((autograd.backproject(yg, g) - c) ** 2).sum()).backward()
    ref2 = A64(AT64(b) - c.double().cpu())
    assert rel_l2(yg.grad, ref2) <= 1e-5


def test_fista_tv_fused_plain_and_oracle():
    """10 FISTA-TV iterations: fused kernel path vs plain torch path (GPU) vs
    the same iteration in FP64 on the oracle; RMSE / peak <= 1e-4."""
    cuda = torch.device("cuda", 0)
    g = configs.fd_32()
    _case("fista_tv fd_32")
    A64, AT64 = _oracle_ops(g)
    ph = phantom.phantom((32, 32, 32), 1.0, 6, seed=9)
    b64 = A64(ph.double())
    lip = unroll.power_norm(g, device=cuda)
    lam, eps, iters = 0.02, 0.01, 10
    step = 1.0 / (lip + lam * 12.0 / eps)
    xf = unroll.fista_tv(b64.float().to(cuda), g, iters, lam, eps, step=step, fused=True)
    xp = unroll.fista_tv(b64.float().to(cuda), g, iters, lam, eps, step=step, fused=False)
    # FP64 oracle run of the same iteration
    x = torch.zeros_like(ph, dtype=torch.float64)
    y = x.clone()
    for beta in unroll.fista_momentum(iters):
        gr = AT64(A64(y) - b64)
        x, y = tv.fista_tv_step_torch(y, gr, x, step, lam, eps, beta)
    peak = float(x.abs().max())
    for got in (xf, xp):
        rmse = float(((got.double().cpu() - x) ** 2).mean().sqrt())
        assert rmse / peak <= 1e-4
    assert float(x.abs().sum()) > 0


def test_fista_64_reconstruction_converges():
    """The 64^3 / 72-view workload: 10 fused FISTA-TV iterations reduce the
    data residual and move towards the phantom."""
    cuda = torch.device("cuda", 0)
    g = configs.fista_64()
    _case("fista_64 recon")
    ph = phantom.phantom((64, 64, 64), 1.0, 10, seed=1).to(cuda)
    b = ops.forward(ph, g)
    x = unroll.fista_tv(b, g, 10, lam=1e-3, eps=1e-3)
    r0 = float(b.norm())
    r1 = float((ops.forward(x, g) - b).norm())
    assert r1 < 0.2 * r0
    assert float((x - ph).norm()) < float(ph.norm())


def test_pdhg_dtv_net_matches_fp64_oracle_network():
    """Forward output and parameter gradients of the checkpointed learned
    unroll on the GPU vs the identical network on FP64 oracle operators."""
    cuda = torch.device("cuda", 0)
    g = configs.fd_32()
    _case("pdhg_dtv fd_32")
    A64, AT64 = _oracle_ops(g)
    ph = phantom.phantom((32, 32, 32), 1.0, 6, seed=3)
    b = A64(ph.double())
    net = unroll.PDHGDTVNet(g, iters=3, norm=unroll.power_norm(g, device=cuda)).to(cuda)
    out = net(b.float().to(cuda))
    loss = ((out - ph.to(cuda)) ** 2).mean()
    loss.backward()
    net64 = unroll.PDHGDTVNet(g, iters=3, norm=net.knorm ** 2 - 12.0, project=A64,
                              backproject=AT64, checkpointing=False).double()
    with torch.no_grad():
        for p64, p in zip(net64.parameters(), net.parameters()):
            p64.copy_(p.double().cpu())
    out64 = net64(b)
    loss64 = ((out64 - ph.double()) ** 2).mean()
    loss64.backward()
    assert rel_l2(out, out64) <= 1e-4
    for p, p64 in zip(net.parameters(), net64.parameters()):
        assert rel_l2(p.grad, p64.grad) <= 1e-3
    # FP64 central differences on the oracle network validate its autograd
    with torch.no_grad():
        v = torch.zeros_like(net64.sigma)
        v[1] = 1.0
        h = 1e-4
        net64.sigma += h * v
        lp = float(((net64(b) - ph.double()) ** 2).mean())
        net64.sigma -= 2 * h * v
        lm = float(((net64(b) - ph.double()) ** 2).mean())
        net64.sigma += h * v
    fd = (lp - lm) / (2 * h)
    assert abs(fd - float(net64.sigma.grad[1])) <= 1e-6 + 1e-4 * abs(fd)


def test_pdhg_dtv_net_fused_inference_matches_torch_iteration():
    """No-grad inference on the fused PDHG launches vs the torch ops of the
    same network (both on the gfx950 projector pair)."""
    cuda = torch.device("cuda", 0)
    g = configs.fd_32()
    _case("pdhg_dtv fused inference fd_32")
    ph = phantom.phantom((32, 32, 32), 1.0, 6, seed=4).to(cuda)
    with torch.no_grad():
        b = ops.forward(torch.cat([ph, 0.5 * ph]), g)
    for mode in ("neumann", "reference"):
        net = unroll.PDHGDTVNet(g, iters=6, norm=unroll.power_norm(g, device=cuda),
                                lam=(2e-2, 1e-2, 3e-2), mode=mode).to(cuda)
        with torch.no_grad():
            fused = net(b)
            net.fused = False
            plain = net(b)
        assert rel_l2(fused, plain) <= 1e-5
        assert float(plain.norm()) > 0


def test_checkpointing_same_gradients_less_memory():
    cuda = torch.device("cuda", 0)
    g = configs.fd_32()
    b = ops.forward(phantom.phantom((32, 32, 32), 1.0, 6, seed=4).to(cuda), g)
    grads = []
    peaks = []
    for ckpt in (False, True):
        torch.manual_seed(0)
        net = unroll.PDHGDTVNet(g, iters=6, checkpointing=ckpt).to(cuda)
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        out = net(b.repeat(2, 1, 1, 1, 1))
        out.square().mean().backward()
        torch.cuda.synchronize()
        peaks.append(torch.cuda.max_memory_allocated() - base)
        grads.append(torch.cat([p.grad.reshape(-1) for p in net.parameters()]))
    assert rel_l2(grads[1], grads[0]) <= 1e-4
    assert peaks[1] < peaks[0]


def test_asd_pocs_matches_fp64_oracle_run():
    """4 ASD-POCS cascades (data step + 10 TV descent steps each): fused
    kernel path and plain torch path on the GPU vs the same cascades on the
    FP64 oracle operators; RMSE / peak <= 1e-4."""
    cuda = torch.device("cuda", 0)
    g = configs.fd_32()
    _case("asd_pocs fd_32")
    A64, AT64 = _oracle_ops(g)
    ph = phantom.phantom((32, 32, 32), 1.0, 6, seed=9)
    b64 = A64(ph.double())
    lip = unroll.power_norm(g, device=cuda)
    kw = dict(cascades=4, beta=1.0 / lip, ng=10, alpha=0.2)
    proj = lambda x: ops.forward(x, g)
    back = lambda y: ops.adjoint(y, g)
    b = b64.float().to(cuda)
    x0 = torch.zeros(g.volume_shape(1), device=cuda)
    fused = unroll.asd_pocs(x0, b, proj, back, **kw)
    plain = unroll.asd_pocs(x0, b, proj, back, fused=False, **kw)
    ref = unroll.asd_pocs(torch.zeros(g.volume_shape(1), dtype=torch.float64), b64, A64, AT64,
                          fused=False, **kw)
    peak = float(ref.abs().max())
    for got in (fused, plain):
        rmse = float(((got.double().cpu() - ref) ** 2).mean().sqrt())
        assert rmse / peak <= 1e-4
    # the TV steps change the result: without them it is a different image
    no_tv = unroll.asd_pocs(x0, b, proj, back, **dict(kw, ng=0))
    assert rel_l2(no_tv, ref) > 1e-3


def test_dtv_cp_matches_fp64_oracle_run():
    """5 directional-TV Chambolle-Pock cascades with the l1-ball dual
    constraints on the GPU vs the same cascades on the FP64 oracle."""
    cuda = torch.device("cuda", 0)
    g = configs.fd_32()
    _case("dtv_cp fd_32")
    A64, AT64 = _oracle_ops(g)
    ph = phantom.phantom((32, 32, 32), 1.0, 6, seed=10)
    b64 = A64(ph.double())
    lip = unroll.power_norm(g, device=cuda)
    start = 0.5 * ph.double()
    kw = dict(cascades=5, v=(0.5, 0.5, 0.5), mu=1.0, sigma=0.5 / lip, tau=0.5 / lip)
    got = unroll.dtv_cp(start.float().to(cuda), b64.float().to(cuda),
                        lambda x: ops.forward(x, g), lambda y: ops.adjoint(y, g), **kw)
    plain = unroll.dtv_cp(start.float().to(cuda), b64.float().to(cuda),
                          lambda x: ops.forward(x, g), lambda y: ops.adjoint(y, g),
                          fused=False, **kw)
    ref = unroll.dtv_cp(start, b64, A64, AT64, fused=False, **kw)
    peak = float(ref.abs().max())
    for x in (got, plain):
        rmse = float(((x.double().cpu() - ref) ** 2).mean().sqrt())
        assert rmse / peak <= 1e-4
    # the iteration moved the data residual down from the start
    r0 = float((A64(start) - b64).norm())
    assert float((A64(ref) - b64).norm()) < r0


def test_fista_l1_matches_fp64_oracle_run():
    """8 cascades of the soft-threshold FISTA of FISTA.py:16-27 (step 1/L of
    the matched pair, threshold lamb L): fused shrink + momentum kernel and
    torch ops on the GPU vs the same cascades on the FP64 oracle operators."""
    cuda = torch.device("cuda", 0)
    g = configs.fd_32()
    _case("fista_l1 fd_32")
    A64, AT64 = _oracle_ops(g)
    ph = phantom.phantom((32, 32, 32), 1.0, 6, seed=12)
    b64 = A64(ph.double())
    lip = unroll.power_norm(g, device=cuda)
    kw = dict(cascades=8, L=1.0 / lip, lamb=0.05 * lip)
    proj = lambda x: ops.forward(x, g)
    back = lambda y: ops.adjoint(y, g)
    x0 = torch.zeros(g.volume_shape(1), device=cuda)
    fused = unroll.fista_l1(x0, b64.float().to(cuda), proj, back, **kw)
    plain = unroll.fista_l1(x0, b64.float().to(cuda), proj, back, fused=False, **kw)
    ref = unroll.fista_l1(torch.zeros(g.volume_shape(1), dtype=torch.float64), b64, A64, AT64,
                          fused=False, **kw)
    peak = float(ref.abs().max())
    assert float((ref == 0).double().mean()) > 0.05      # the threshold acts
    for x in (fused, plain):
        rmse = float(((x.double().cpu() - ref) ** 2).mean().sqrt())
        assert rmse / peak <= 1e-4
    assert float(x0.abs().max()) == 0.0                   # the start is not modified


def test_fdpg_dtv_matches_fp64_oracle_run():
    """6 cascades of the directional-TV dual-gradient iteration of
    DTVFISTA.py:23-43 (reference-mode difference matrices): fused step and
    torch ops on the GPU vs the same cascades on the FP64 oracle operators."""
    cuda = torch.device("cuda", 0)
    g = configs.fd_32()
    _case("fdpg_dtv fd_32")
    A64, AT64 = _oracle_ops(g)
    ph = phantom.phantom((32, 32, 32), 1.0, 6, seed=13)
    b64 = A64(ph.double())
    lip = unroll.power_norm(g, device=cuda)
    kw = dict(cascades=6, sigma=1.0 / lip, v=(0.02, 0.03, 0.02), t=(0.05, 0.05, 0.04))
    proj = lambda x: ops.forward(x, g)
    back = lambda y: ops.adjoint(y, g)
    # a noisy start: the iteration takes sign(D_a z), which FP32 and FP64
    # would only disagree on where D_a z rounds to about zero -- with noise
    # of 0.05 almost nowhere
    start = 0.5 * ph + 0.05 * phantom.gaussian(ph.shape, 14)
    fused = unroll.fdpg_dtv(start.to(cuda), b64.float().to(cuda), proj, back, **kw)
    plain = unroll.fdpg_dtv(start.to(cuda), b64.float().to(cuda), proj, back, fused=False, **kw)
    ref = unroll.fdpg_dtv(start.double(), b64, A64, AT64, fused=False, **kw)
    peak = float(ref.abs().max())
    for x in (fused, plain):
        rmse = float(((x.double().cpu() - ref) ** 2).mean().sqrt())
        assert rmse / peak <= 1e-4
    # the TV part moves the result away from the plain projected gradient step
    no_tv = unroll.fdpg_dtv(start.to(cuda), b64.float().to(cuda), proj, back,
                            **dict(kw, t=(0.0, 0.0, 0.0)))
    assert rel_l2(no_tv, ref) > 1e-4
